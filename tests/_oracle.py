"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test
infrastructure only: it is the checker for the GPU library, never the thing
measured or shipped."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

EVEN, ODD, EVENODD = 0, 1, 2


class Coeffs(C.Structure):
    _fields_ = [("one_link", C.c_double), ("three_staple", C.c_double),
                ("five_staple", C.c_double), ("seven_staple", C.c_double),
                ("lepage", C.c_double), ("naik", C.c_double)]


class InvArg(C.Structure):
    _fields_ = [("mixed_rsq", C.c_double), ("max_iter", C.c_int), ("restart", C.c_int),
                ("max_restarts", C.c_int), ("evenodd", C.c_int)]


class ResArg(C.Structure):
    _fields_ = [("rsqmin", C.c_double), ("final_rsq", C.c_double), ("relmin", C.c_double),
                ("final_rel", C.c_double), ("final_iter", C.c_int), ("final_restart", C.c_int)]


class ReunitOpts(C.Structure):
    _fields_ = [("allow_svd", C.c_int), ("svd_only", C.c_int), ("svd_rel_error", C.c_double),
                ("svd_abs_error", C.c_double)]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_PP = C.POINTER(C.c_void_p)
_IV = C.POINTER(C.c_int)


def _iv(x):
    return (C.c_int * len(x))(*x)


def _pa(arrs):
    a = (C.c_void_p * len(arrs))()
    for i, x in enumerate(arrs):
        a[i] = x.ctypes.data if x is not None else None
    return a


class Oracle:
    _lib = None

    def __init__(self):
        if Oracle._lib is None:
            if not os.path.exists(ORACLE_SO):
                raise RuntimeError("oracle/liboracle.so not built (run __graft_entry__.build())")
            L = C.CDLL(ORACLE_SO)
            L.or_volume.restype = C.c_long
            L.or_volume.argtypes = [_IV]
            L.or_raw_index.restype = C.c_long
            L.or_raw_index.argtypes = [_IV, _IV]
            L.or_raw_coords.argtypes = [_IV, C.c_long, _IV]
            L.or_rephase_G.argtypes = [_IV, _PP, _IV, C.POINTER(C.c_double), _IV]
            L.or_make_imp_links.argtypes = [_IV, C.POINTER(Coeffs), _PP, _PP, _PP, _PP]
            L.or_plaquette.restype = C.c_double
            L.or_plaquette.argtypes = [_IV, _PP]
            L.or_links_create.restype = C.c_void_p
            L.or_links_create.argtypes = [_IV, _PP, _PP]
            L.or_links_destroy.argtypes = [C.c_void_p]
            L.or_dslash.argtypes = [C.c_void_p, C.c_double, _D, _D, C.c_int, C.c_int]
            L.or_dslash_check_op.argtypes = [_IV, _PP, _PP, C.c_double, _D, _D]
            L.or_dslash_dir.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, _D, _D,
                                        C.c_int, C.c_double]
            L.or_links_set_single.argtypes = [C.c_void_p, C.c_int]
            L.or_ddagd_norm2.restype = C.c_double
            L.or_ddagd_norm2.argtypes = [C.c_void_p, C.c_double, _D, _D, C.c_int]
            L.or_invert.argtypes = [C.c_void_p, C.POINTER(InvArg), C.POINTER(ResArg), C.c_double,
                                    _D, _D, C.POINTER(C.c_double)]
            L.or_invert_multi.argtypes = [C.c_void_p, C.POINTER(InvArg), C.POINTER(ResArg),
                                          _D, _IV, _PP, _PP, C.c_int, C.POINTER(C.c_double)]
            L.or_eigcg_create.restype = C.c_void_p
            L.or_eigcg_destroy.argtypes = [C.c_void_p]
            L.or_eigcg_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
            L.or_eigcg_nu.restype = C.c_int
            L.or_eigcg_nu.argtypes = [C.c_void_p, _D, C.c_int]
            L.or_invert_eigcg.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(InvArg), C.POINTER(ResArg),
                                          C.c_double, _D, _D, C.POINTER(C.c_double)]
            L.or_relnorm2.restype = C.c_double
            L.or_relnorm2.argtypes = [C.c_void_p, _D, _D, C.c_int]
            L.or_u3_un_analytic.restype = C.c_int
            L.or_u3_un_analytic.argtypes = [_D, _D, C.POINTER(ReunitOpts), C.POINTER(C.c_double)]
            L.or_u3_polar_jacobi.argtypes = [_D, _D]
            L.or_u3_deriv_contract.restype = C.c_int
            L.or_u3_deriv_contract.argtypes = [_D, _D, C.c_double, _D]
            L.or_hisq_force_multi.restype = C.c_int
            L.or_hisq_force_multi.argtypes = [_IV, _PP, _PP, C.POINTER(Coeffs), C.POINTER(Coeffs),
                                              C.c_double, C.c_double, C.c_int,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double), _PP,
                                              C.POINTER(C.c_int), C.c_double]
            L.or_hisq_W.restype = C.c_int
            L.or_hisq_W.argtypes = [_IV, C.POINTER(Coeffs), _PP, _PP, C.POINTER(ReunitOpts),
                                    C.POINTER(C.c_double)]
            L.or_asqtad_force_multi.argtypes = [_IV, _PP, _PP, C.POINTER(Coeffs),
                                                C.POINTER(C.c_double), _PP, C.c_int]
            L.or_set_verbose.restype = C.c_int
            L.or_num_threads.restype = C.c_int
            Oracle._lib = L
        self.lib = Oracle._lib

    # geometry
    def raw_index(self, L, x):
        return self.lib.or_raw_index(_iv(L), _iv(x))

    def raw_coords(self, L, i):
        x = (C.c_int * 4)()
        self.lib.or_raw_coords(_iv(L), i, x)
        return list(x)

    def rephase(self, L, U, r0, bc, signmask):
        bcarr = (C.c_double * 8)(*[v for c in bc for v in (c.real, c.imag)])
        self.lib.or_rephase_G(_iv(L), _pa(U), _iv(r0), bcarr, _iv(signmask) if signmask else None)

    def rephase_milc(self, L, U):
        self.rephase(L, U, [0, 0, 0, 0], [1, 1, 1, -1], [0, 1, 3, 7])

    def make_imp_links(self, L, U, coeffs, Ulong=None, with_long=None):
        V = int(np.prod(L))
        c = Coeffs(**coeffs)
        fat = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        if with_long is None:
            with_long = coeffs["naik"] != 0
        lng = [np.zeros((V, 3, 3, 2)) for _ in range(4)] if with_long else None
        self.lib.or_make_imp_links(_iv(L), C.byref(c), _pa(U), _pa(Ulong if Ulong else U),
                                   _pa(fat), _pa(lng) if lng else None)
        return fat, lng

    def asqtad_force(self, L, U, coeffs, eps, xs, force=None):
        """force[mu] += -2 antiherm(U deriv^+) for sum_i eps_i x_i x_i^+
        (lib/asqtad_force_fnmat_p.c:515-546); xs are full-lattice vectors."""
        V = int(np.prod(L))
        if force is None:
            force = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        self.lib.or_asqtad_force_multi(_iv(L), _pa(U), _pa(force), C.byref(Coeffs(**coeffs)),
                                       (C.c_double * len(eps))(*eps), _pa(xs), len(xs))
        return force

    def plaquette(self, L, U):
        return self.lib.or_plaquette(_iv(L), _pa(U))

    def links(self, L, fat, lng):
        fat = [np.ascontiguousarray(f) for f in fat]
        lng = [np.ascontiguousarray(x) for x in lng] if lng is not None else None
        return self.lib.or_links_create(_iv(L), _pa(fat), _pa(lng) if lng else None)

    def free_links(self, h):
        self.lib.or_links_destroy(h)

    def dslash(self, h, mass, vin, eo_out, eo_in, out=None):
        vin = np.ascontiguousarray(vin, dtype=np.float64)
        if out is None:
            out = np.zeros_like(vin)
        self.lib.or_dslash(h, mass, out.reshape(-1), vin.reshape(-1), eo_out, eo_in)
        return out

    def dslash_dir(self, h, dir, fb, wtfat, wtlong, vin, eo_out, out, bsign=1.0):
        """out += one-direction transport (QOP_asqtad_dslash_dir); out is updated in place."""
        vin = np.ascontiguousarray(vin, dtype=np.float64)
        self.lib.or_dslash_dir(h, dir, fb, wtfat, wtlong, out.reshape(-1), vin.reshape(-1), eo_out, bsign)
        return out

    def set_single(self, h, on=True):
        """FP32 storage of the links and of every stored vector (the F3 build)."""
        self.lib.or_links_set_single(h, 1 if on else 0)

    def check_op(self, L, fat, lng, mass, vin):
        out = np.zeros_like(vin)
        self.lib.or_dslash_check_op(_iv(L), _pa(fat), _pa(lng) if lng else None, mass,
                                    out.reshape(-1), np.ascontiguousarray(vin).reshape(-1))
        return out

    def ddagd_norm2(self, h, mass, vin, eo):
        out = np.zeros_like(vin)
        r = self.lib.or_ddagd_norm2(h, mass, out.reshape(-1), np.ascontiguousarray(vin).reshape(-1), eo)
        return out, r

    def invert(self, h, mass, vin, evenodd, rsqmin, out=None, max_iter=2000, restart=1000,
               max_restarts=5, relmin=0.0, mixed_rsq=1e9):
        ia = InvArg(mixed_rsq, max_iter, restart, max_restarts, evenodd)
        ra = ResArg(rsqmin, 0, relmin, 0, 0, 0)
        if out is None:
            out = np.zeros_like(vin)
        flop = C.c_double(0)
        self.lib.or_invert(h, C.byref(ia), C.byref(ra), mass, out.reshape(-1),
                           np.ascontiguousarray(vin).reshape(-1), C.byref(flop))
        return out, ra, flop.value

    def eigcg_state(self):
        return self.lib.or_eigcg_create()

    def eigcg_free(self, eg):
        self.lib.or_eigcg_destroy(eg)

    def eigcg_ritz(self, eg, maxn=1000):
        l = np.zeros(maxn)
        n = self.lib.or_eigcg_nu(eg, l, maxn)
        return l[:min(n, maxn)]

    def invert_eigcg(self, h, eg, mass, vin, rsqmin, numax, m, nev, out=None, max_iter=2000,
                     restart=1000, max_restarts=5):
        """EVENODD QOP_asqtad_invert with cgtype = 1 (eigCG), state eg kept across calls."""
        self.lib.or_eigcg_set(eg, numax, m, nev)
        ia = InvArg(1e9, max_iter, restart, max_restarts, EVENODD)
        ra = ResArg(rsqmin, 0, 0, 0, 0, 0)
        if out is None:
            out = np.zeros_like(vin)
        flop = C.c_double(0)
        self.lib.or_invert_eigcg(h, eg, C.byref(ia), C.byref(ra), mass, out.reshape(-1),
                                 np.ascontiguousarray(vin).reshape(-1), C.byref(flop))
        return out, ra

    def invert_multi(self, h, masses, vins, evenodd, rsqmin, max_iter=2000, restart=1000,
                     max_restarts=5, relmin=0.0, mixed_rsq=1e9):
        """masses: list (per source) of lists of masses."""
        ia = InvArg(mixed_rsq, max_iter, restart, max_restarts, evenodd)
        nmass = [len(m) for m in masses]
        ns = sum(nmass)
        ra = (ResArg * ns)(*[ResArg(rsqmin, 0, relmin, 0, 0, 0) for _ in range(ns)])
        flat = np.array([m for ms in masses for m in ms], dtype=np.float64)
        outs = [np.zeros_like(vins[0]) for _ in range(ns)]
        ins = [np.ascontiguousarray(v) for v in vins]
        flop = C.c_double(0)
        self.lib.or_invert_multi(h, C.byref(ia), ra, flat, _iv(nmass), _pa(outs), _pa(ins),
                                 len(vins), C.byref(flop))
        return outs, list(ra), flop.value

    def relnorm2(self, h, r, x, eo):
        return self.lib.or_relnorm2(h, np.ascontiguousarray(r).reshape(-1),
                                    np.ascontiguousarray(x).reshape(-1), eo)

    # derived checks
    def true_resid_even(self, h, mass, out, src, eo=EVEN):
        """|src - A out/m| / |src| on subset eo with A = m^2 - D D
        (the check_resids identity, examples/asqtad_test.c:84-103)."""
        x = out / mass
        Ax, _ = self.ddagd_norm2(h, mass, x, eo)
        V = src.shape[0]
        sl = slice(0, V // 2) if eo == EVEN else slice(V // 2, V)
        r = src[sl] - Ax[sl]
        return float(np.linalg.norm(r) / np.linalg.norm(src[sl]))

    def true_resid_full(self, h, mass, out, src):
        """|src - D out| / |src| over all sites."""
        Dx = self.dslash(h, mass, out, EVENODD, EVENODD)
        return float(np.linalg.norm(src - Dx) / np.linalg.norm(src))

    # HISQ preparation
    @staticmethod
    def reunit_opts(allow_svd=1, svd_only=0, rel=1e-8, abs_=1e-8):
        return ReunitOpts(allow_svd, svd_only, rel, abs_)

    def u3_un_analytic(self, V, opts=None):
        """V: (3,3) complex -> (W, used_fallback, flops)"""
        o = opts or self.reunit_opts()
        v = np.ascontiguousarray(np.stack([V.real, V.imag], -1), dtype=np.float64).reshape(-1)
        w = np.zeros(18)
        fl = C.c_double(0)
        svd = self.lib.or_u3_un_analytic(v, w, C.byref(o), C.byref(fl))
        w = w.reshape(3, 3, 2)
        return w[..., 0] + 1j * w[..., 1], svd, fl.value

    def u3_polar_jacobi(self, V):
        v = np.ascontiguousarray(np.stack([V.real, V.imag], -1), dtype=np.float64).reshape(-1)
        w = np.zeros(18)
        self.lib.or_u3_polar_jacobi(v, w)
        w = w.reshape(3, 3, 2)
        return w[..., 0] + 1j * w[..., 1]

    def u3_deriv_contract(self, V, F, filt=0.0):
        """Y(n,m) = d[tr(W F^+) + tr(W^+ F)]/dV_mn for (3,3) complex V, F; (Y, filtered)"""
        v = np.ascontiguousarray(np.stack([V.real, V.imag], -1), dtype=np.float64).reshape(-1)
        f = np.ascontiguousarray(np.stack([F.real, F.imag], -1), dtype=np.float64).reshape(-1)
        y = np.zeros(18)
        k = self.lib.or_u3_deriv_contract(v, f, filt, y)
        y = y.reshape(3, 3, 2)
        return y[..., 0] + 1j * y[..., 1], k

    def hisq_force(self, L, U, fat7, lev2, diff_one_link, diff_naik, eps_naik, eps, xs, n_orders,
                   force_filter=0.0, force=None):
        """force[mu] += antiherm(U deriv^+) of the HISQ fnmat force
        (lib/hisq_force_fnmat_p.c:7-231); returns (force, filtered links)."""
        V = int(np.prod(L))
        if force is None:
            force = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        n = len(n_orders)
        k = self.lib.or_hisq_force_multi(
            _iv(L), _pa([np.ascontiguousarray(u) for u in U]), _pa(force), C.byref(Coeffs(**fat7)),
            C.byref(Coeffs(**lev2)), diff_one_link, diff_naik, n, (C.c_double * n)(*eps_naik),
            (C.c_double * len(eps))(*eps), _pa(xs), (C.c_int * n)(*n_orders), force_filter)
        return force, k

    def hisq_W(self, L, U, fat7, opts=None):
        """fat7 smearing + U(3) projection: list of 4 (V,3,3,2) arrays, #fallbacks, flops"""
        V = int(np.prod(L))
        W = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
        o = opts or self.reunit_opts()
        c = Coeffs(**fat7)
        fl = C.c_double(0)
        n = self.lib.or_hisq_W(_iv(L), C.byref(c), _pa([np.ascontiguousarray(u) for u in U]), _pa(W),
                               C.byref(o), C.byref(fl))
        return W, n, fl.value

    def threads(self):
        return self.lib.or_num_threads()
