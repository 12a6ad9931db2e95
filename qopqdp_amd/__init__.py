"""Python binding of the MI355X asqtad library (libqop_mi355x.so).

The library exports the QOP C API of the asqtad hot path (see
include/qop_mi355x.h).  This module binds it with ctypes and offers small
helpers that move numpy arrays in the QOP raw layout in and out, so tests and
benchmarks read like the reference's own example programs
(examples/asqtad_test.c, examples/asqtad_invert_bench.c).

There is no CPU fallback: if the shared library is missing, or no HIP device
is present when a compute entry point is called, an error is raised.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBNAME = os.path.join(_HERE, "libqop_mi355x.so")

QOP_EVEN, QOP_ODD, QOP_EVENODD = 0, 1, 2
QOP_SUCCESS, QOP_FAIL = 0, 1


class QOPInfo(C.Structure):
    _fields_ = [("final_sec", C.c_double), ("final_flop", C.c_double),
                ("status", C.c_int), ("count1", C.c_int), ("count2", C.c_int)]


class QOPInvertArg(C.Structure):
    _fields_ = [("mixed_rsq", C.c_double), ("max_iter", C.c_int), ("restart", C.c_int),
                ("max_restarts", C.c_int), ("evenodd", C.c_int)]


class QOPResidArg(C.Structure):
    _fields_ = [("rsqmin", C.c_double), ("final_rsq", C.c_double), ("relmin", C.c_double),
                ("final_rel", C.c_double), ("final_iter", C.c_int), ("final_restart", C.c_int)]


class QOPLayout(C.Structure):
    _fields_ = [("node_number", C.c_void_p), ("node_index", C.c_void_p), ("latdim", C.c_int),
                ("latsize", C.POINTER(C.c_int)), ("machdim", C.c_int),
                ("machsize", C.POINTER(C.c_int)), ("this_node", C.c_int),
                ("sites_on_node", C.c_int)]


class QOPComplex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class QOPBc(C.Structure):
    _fields_ = [("phase", C.POINTER(QOPComplex))]


class QOPSign(C.Structure):
    _fields_ = [("signmask", C.POINTER(C.c_int))]


class QOPOpt(C.Structure):
    _fields_ = [("tag", C.c_char_p), ("value", C.c_double), ("extra", C.c_void_p)]


class QOPHisqCoeffs(C.Structure):
    _fields_ = [("n_naiks", C.c_int), ("eps_naik", C.c_double * 20), ("ugroup", C.c_int),
                ("umethod", C.c_int)] + [(n, C.c_double) for n in (
                    "fat7_one_link", "fat7_three_staple", "fat7_five_staple", "fat7_seven_staple",
                    "fat7_lepage", "asqtad_one_link", "asqtad_three_staple", "asqtad_five_staple",
                    "asqtad_seven_staple", "asqtad_lepage", "asqtad_naik", "difference_one_link",
                    "difference_naik")]


# MILC HISQ coefficients (fat7 level, then the asqtad level 2 with
# epsilon_Naik = 0); free-field fat weights 1 and 9/8
MILC_HISQ = dict(fat7_one_link=1 / 8, fat7_three_staple=-1 / 16, fat7_five_staple=1 / 64,
                 fat7_seven_staple=-1 / 384, fat7_lepage=0.0, asqtad_one_link=1.0,
                 asqtad_three_staple=-1 / 16, asqtad_five_staple=1 / 64,
                 asqtad_seven_staple=-1 / 384, asqtad_lepage=-1 / 8, asqtad_naik=-1 / 24,
                 difference_one_link=0.0, difference_naik=-1 / 24)


def hisq_coeffs(eps_naik=(0.0,), **kw):
    c = QOPHisqCoeffs()
    c.n_naiks = len(eps_naik)
    for i, e in enumerate(eps_naik):
        c.eps_naik[i] = e
    c.ugroup, c.umethod = 0, 3
    for k, v in {**MILC_HISQ, **kw}.items():
        setattr(c, k, v)
    return c


class QOPAsqtadCoeffs(C.Structure):
    _fields_ = [("one_link", C.c_double), ("three_staple", C.c_double),
                ("five_staple", C.c_double), ("seven_staple", C.c_double),
                ("lepage", C.c_double), ("naik", C.c_double)]


def invert_arg_default():
    """QOP_INVERT_ARG_DEFAULT (include/qop_int.h:126)."""
    return QOPInvertArg(1e9, 2000, 1000, 5, QOP_EVENODD)


def resid_arg_default():
    """QOP_RESID_ARG_DEFAULT (include/qop_int.h:137)."""
    return QOPResidArg(1e-6, 0, 0, 0, 0, 0)


MILC_ASQTAD = dict(one_link=5.0 / 8, three_staple=-1.0 / 16, five_staple=1.0 / 64,
                   seven_staple=-1.0 / 384, lepage=-1.0 / 16, naik=-1.0 / 24)

_lib = None

_VP = C.c_void_p
_DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_FP = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_PP = C.POINTER(C.c_void_p)

# (name, restype, argtypes) for every exported symbol of include/qop_mi355x.h
SIGNATURES = [
    ("QOP_init", C.c_int, [C.POINTER(QOPLayout)]),
    ("QOP_finalize", C.c_int, []),
    ("QOP_is_initialized", C.c_int, []),
    ("QOP_version_str", C.c_char_p, []),
    ("QOP_version_int", C.c_int, []),
    ("QOP_verbose", C.c_int, [C.c_int]),
    ("QOP_profcontrol", C.c_int, [C.c_int]),
    ("QOP_node_number_raw", C.c_int, [C.POINTER(C.c_int)]),
    ("QOP_node_index_raw_V", C.c_int, [C.POINTER(C.c_int), C.c_int]),
    ("QOP_node_index_raw_G", C.c_int, [C.POINTER(C.c_int), C.c_int]),
    ("QOP_sites_on_node_raw_V", C.c_int, [C.c_int]),
    ("QOP_sites_on_node_raw_G", C.c_int, [C.c_int]),
    ("QOP_asqtad_invert_set_opts", C.c_int, [C.POINTER(QOPOpt), C.c_int]),
    ("QOP_D3_create_V_from_raw", _VP, [_DP, C.c_int]),
    ("QOP_F3_create_V_from_raw", _VP, [_FP, C.c_int]),
    ("QOP_D3_extract_V_to_raw", None, [_DP, _VP, C.c_int]),
    ("QOP_F3_extract_V_to_raw", None, [_FP, _VP, C.c_int]),
    ("QOP_D3_destroy_V", None, [_VP]),
    ("QOP_F3_destroy_V", None, [_VP]),
    ("QOP_D3_convert_V_from_raw", _VP, [_DP, C.c_int]),
    ("QOP_F3_convert_V_from_raw", _VP, [_FP, C.c_int]),
    ("QOP_D3_convert_V_to_raw", C.c_void_p, [_VP, C.c_int]),
    ("QOP_F3_convert_V_to_raw", C.c_void_p, [_VP, C.c_int]),
    ("QOP_D3_create_G_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_F3_create_G_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_D3_extract_G_to_raw", None, [_PP, _VP, C.c_int]),
    ("QOP_F3_extract_G_to_raw", None, [_PP, _VP, C.c_int]),
    ("QOP_D3_destroy_G", None, [_VP]),
    ("QOP_F3_destroy_G", None, [_VP]),
    ("QOP_D3_convert_G_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_F3_convert_G_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_D3_convert_G_to_raw", C.c_void_p, [_VP, C.c_int]),
    ("QOP_F3_convert_G_to_raw", C.c_void_p, [_VP, C.c_int]),
    ("QOP_D3_rephase_G", None, [_VP, C.POINTER(C.c_int), C.POINTER(QOPBc), C.POINTER(QOPSign)]),
    ("QOP_F3_rephase_G", None, [_VP, C.POINTER(C.c_int), C.POINTER(QOPBc), C.POINTER(QOPSign)]),
    ("QOP_D3_create_F_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_F3_create_F_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_D3_extract_F_to_raw", None, [_PP, _VP, C.c_int]),
    ("QOP_F3_extract_F_to_raw", None, [_PP, _VP, C.c_int]),
    ("QOP_D3_destroy_F", None, [_VP]),
    ("QOP_F3_destroy_F", None, [_VP]),
    ("QOP_D3_convert_F_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_F3_convert_F_from_raw", _VP, [_PP, C.c_int]),
    ("QOP_D3_convert_F_to_raw", C.c_void_p, [_VP, C.c_int]),
    ("QOP_F3_convert_F_to_raw", C.c_void_p, [_VP, C.c_int]),
    ("QOP_asqtad_force_set_opts", C.c_int, [C.POINTER(QOPOpt), C.c_int]),
    ("QOP_D3_asqtad_force", None, [C.POINTER(QOPInfo), _VP, _VP, C.POINTER(QOPAsqtadCoeffs), C.c_double, _VP]),
    ("QOP_F3_asqtad_force", None, [C.POINTER(QOPInfo), _VP, _VP, C.POINTER(QOPAsqtadCoeffs), C.c_float, _VP]),
    ("QOP_D3_asqtad_force_multi", None, [C.POINTER(QOPInfo), _VP, _VP, C.POINTER(QOPAsqtadCoeffs),
                                         C.POINTER(C.c_double), _PP, C.c_int]),
    ("QOP_F3_asqtad_force_multi", None, [C.POINTER(QOPInfo), _VP, _VP, C.POINTER(QOPAsqtadCoeffs),
                                         C.POINTER(C.c_float), _PP, C.c_int]),
    ("QOP_D3_asqtad_create_L_from_raw", _VP, [_PP, _PP, C.c_int]),
    ("QOP_F3_asqtad_create_L_from_raw", _VP, [_PP, _PP, C.c_int]),
    ("QOP_D3_asqtad_create_L_from_G", _VP, [C.POINTER(QOPInfo), C.POINTER(QOPAsqtadCoeffs), _VP]),
    ("QOP_F3_asqtad_create_L_from_G", _VP, [C.POINTER(QOPInfo), C.POINTER(QOPAsqtadCoeffs), _VP]),
    ("QOP_D3_asqtad_create_L_from_G2", _VP, [C.POINTER(QOPInfo), C.POINTER(QOPAsqtadCoeffs), _VP, _VP]),
    ("QOP_F3_asqtad_create_L_from_G2", _VP, [C.POINTER(QOPInfo), C.POINTER(QOPAsqtadCoeffs), _VP, _VP]),
    ("QOP_D3_asqtad_extract_L_to_raw", None, [_PP, _PP, _VP, C.c_int]),
    ("QOP_F3_asqtad_extract_L_to_raw", None, [_PP, _PP, _VP, C.c_int]),
    ("QOP_D3_asqtad_convert_L_from_raw", _VP, [_PP, _PP, C.c_int]),
    ("QOP_F3_asqtad_convert_L_from_raw", _VP, [_PP, _PP, C.c_int]),
    ("QOP_D3_asqtad_convert_L_to_raw", None, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _VP, C.c_int]),
    ("QOP_F3_asqtad_convert_L_to_raw", None, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _VP, C.c_int]),
    ("QOP_D3_asqtad_load_L_from_raw", None, [_VP, _PP, _PP, C.c_int]),
    ("QOP_F3_asqtad_load_L_from_raw", None, [_VP, _PP, _PP, C.c_int]),
    ("QOP_D3_asqtad_load_L_from_G", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPAsqtadCoeffs), _VP]),
    ("QOP_F3_asqtad_load_L_from_G", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPAsqtadCoeffs), _VP]),
    ("QOP_D3_asqtad_load_L_from_G2", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPAsqtadCoeffs), _VP, _VP]),
    ("QOP_F3_asqtad_load_L_from_G2", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPAsqtadCoeffs), _VP, _VP]),
    ("QOP_D3_asqtad_destroy_L", None, [_VP]),
    ("QOP_F3_asqtad_destroy_L", None, [_VP]),
    ("QOP_D3_asqtad_rephase_L", None, [_VP, C.POINTER(C.c_int), C.POINTER(QOPBc), C.POINTER(QOPSign)]),
    ("QOP_F3_asqtad_rephase_L", None, [_VP, C.POINTER(C.c_int), C.POINTER(QOPBc), C.POINTER(QOPSign)]),
    ("QOP_hisq_links_set_opts", C.c_int, [C.POINTER(QOPOpt), C.c_int]),
    ("QOP_D3_hisq_create_L_from_G", _VP, [C.POINTER(QOPInfo), C.POINTER(QOPHisqCoeffs), _VP]),
    ("QOP_F3_hisq_create_L_from_G", _VP, [C.POINTER(QOPInfo), C.POINTER(QOPHisqCoeffs), _VP]),
    ("QOP_hisq_force_set_opts", C.c_int, [C.POINTER(QOPOpt), C.c_int]),
    ("QOP_D3_hisq_force_multi", None, [C.POINTER(QOPInfo), _VP, _VP, C.POINTER(QOPHisqCoeffs),
                                       C.POINTER(C.c_double), _PP, C.POINTER(C.c_int)]),
    ("QOP_F3_hisq_force_multi", None, [C.POINTER(QOPInfo), _VP, _VP, C.POINTER(QOPHisqCoeffs),
                                       C.POINTER(C.c_float), _PP, C.POINTER(C.c_int)]),
    ("QOP_D3_hisq_destroy_L", None, [_VP]),
    ("QOP_F3_hisq_destroy_L", None, [_VP]),
    ("QOP_D3_get_asqtad_links_from_hisq", C.POINTER(_VP), [_VP]),
    ("QOP_F3_get_asqtad_links_from_hisq", C.POINTER(_VP), [_VP]),
    ("QOP_D3_get_asqtad_deps_links_from_hisq", _VP, [_VP]),
    ("QOP_F3_get_asqtad_deps_links_from_hisq", _VP, [_VP]),
    ("QOP_FD3_asqtad_create_L_from_L", _VP, [_VP]),
    ("QOP_DF3_asqtad_create_L_from_L", _VP, [_VP]),
    ("QOP_D3_asqtad_dslash", None, [C.POINTER(QOPInfo), _VP, C.c_double, _VP, _VP, C.c_int, C.c_int]),
    ("QOP_F3_asqtad_dslash", None, [C.POINTER(QOPInfo), _VP, C.c_float, _VP, _VP, C.c_int, C.c_int]),
    ("QOP_D3_asqtad_dslash_dir", None, [C.POINTER(QOPInfo), _VP, C.c_int, C.c_int, C.c_double, C.c_double,
                                        _VP, _VP, C.c_int]),
    ("QOP_F3_asqtad_dslash_dir", None, [C.POINTER(QOPInfo), _VP, C.c_int, C.c_int, C.c_double, C.c_double,
                                        _VP, _VP, C.c_int]),
    ("QOP_D3_asqtad_diaginv", None, [C.POINTER(QOPInfo), _VP, C.c_double, _VP, _VP, C.c_int]),
    ("QOP_F3_asqtad_diaginv", None, [C.POINTER(QOPInfo), _VP, C.c_float, _VP, _VP, C.c_int]),
    ("QOP_D3_asqtad_invert", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPInvertArg),
                                    C.POINTER(QOPResidArg), C.c_double, _VP, _VP]),
    ("QOP_F3_asqtad_invert", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPInvertArg),
                                    C.POINTER(QOPResidArg), C.c_float, _VP, _VP]),
    ("QOP_D3_asqtad_invert_multi", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPInvertArg), _PP, _PP,
                                          C.POINTER(C.c_int), _PP, _PP, C.c_int]),
    ("QOP_F3_asqtad_invert_multi", None, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPInvertArg), _PP, _PP,
                                          C.POINTER(C.c_int), _PP, _PP, C.c_int]),
    ("QOPX_gen_gauge", None, [C.POINTER(C.c_int), C.c_ulonglong, C.c_double, _PP]),
    ("QOPX_gen_gaussian_V", None, [C.POINTER(C.c_int), C.c_ulonglong, C.c_int, _DP]),
    ("QOPX_D3_asqtad_invert_batch", C.c_int, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPInvertArg), _PP,
                                              C.POINTER(C.c_double), _PP, _PP, C.c_int, C.c_int, C.c_int,
                                              C.POINTER(C.c_double)]),
    ("QOPX_D3_asqtad_invert_multi_batch", C.c_int, [C.POINTER(QOPInfo), _VP, C.POINTER(QOPInvertArg), _PP,
                                                    C.POINTER(C.c_double), C.c_int, _PP, _PP, C.c_int,
                                                    C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int)]),
    ("QOPX_D3_asqtad_hop", C.c_int, [_VP, C.c_int, C.c_int, _PP, _PP, C.c_int]),
    ("QOPX_D3_asqtad_hop_noalloc", C.c_int, [_VP, C.c_int, C.c_int, _PP, _PP, C.c_int]),
    ("QOPX_D3_asqtad_hop_time", C.c_double, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    ("QOPX_copy_time", C.c_double, [C.c_ulonglong, C.c_int]),
    ("QOPX_hop_shape_override", None, [C.POINTER(C.c_int)]),
    ("QOPX_serve_create", _VP, [C.c_int, C.c_int, C.c_ulonglong]),
    ("QOPX_serve_add_gauge", C.c_int, [_VP, C.POINTER(C.c_int), _PP, C.POINTER(QOPAsqtadCoeffs),
                                       C.POINTER(C.c_int), C.POINTER(QOPBc), C.POINTER(QOPSign)]),
    ("QOPX_serve_add_links", C.c_int, [_VP, C.POINTER(C.c_int), _PP, _PP]),
    ("QOPX_serve_submit", C.c_long, [_VP, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _VP, _VP]),
    ("QOPX_serve_wait", C.c_int, [_VP, C.c_long, C.POINTER(QOPResidArg), C.POINTER(C.c_double)]),
    ("QOPX_serve_test", C.c_int, [_VP, C.c_long]),
    ("QOPX_serve_stats", None, [_VP, C.POINTER(C.c_double)]),
    ("QOPX_serve_destroy", None, [_VP]),
    ("QOPX_device_ok", C.c_int, []),
    ("QOPX_host_alloc", C.c_void_p, [C.c_ulonglong]),
    ("QOPX_host_free", None, [C.c_void_p]),
    ("QOPX_last_error", C.c_char_p, []),
]


def lib():
    """Load libqop_mi355x.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBNAME):
            raise RuntimeError(f"{LIBNAME} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIBNAME)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr_array(arrays):
    """void* array from a list of numpy arrays (kept alive by the caller)."""
    arr = (C.c_void_p * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = a.ctypes.data if a is not None else None
    return arr


def ivec(vals):
    return (C.c_int * len(vals))(*vals)


def volume(L):
    return int(np.prod(L))


def gen_gauge(L, seed, eps=0.3):
    """Seeded SU(3) links U = reunit(exp(i eps H)) in the raw layout: list of 4
    arrays of shape (V, 3, 3, 2)."""
    V = volume(L)
    U = [np.zeros((V, 3, 3, 2)) for _ in range(4)]
    lib().QOPX_gen_gauge(ivec(L), seed, eps, ptr_array(U))
    return U


def gen_source(L, seed, evenodd=QOP_EVEN):
    """Unit-variance complex Gaussian colour vector (V, 3, 2) on `evenodd`."""
    v = np.zeros((volume(L), 3, 2))
    lib().QOPX_gen_gaussian_V(ivec(L), seed, evenodd, v)
    return v


def raw_index(L, x):
    """QOP raw site index (lib/layout_user.c:237-258)."""
    l = ((x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3]
    V = volume(L)
    return (l + V) // 2 if (sum(x) & 1) else l // 2


def raw_coords_all(L):
    """Coordinates of every raw site index, shape (V, 4)."""
    V = volume(L)
    idx = np.arange(V)
    p = (idx >= V // 2).astype(np.int64)
    h = idx - p * (V // 2)

    def lex(l):
        x3 = l % L[3]
        t = l // L[3]
        x2 = t % L[2]
        t //= L[2]
        return np.stack([t // L[1], t % L[1], x2, x3], axis=1)

    x = lex(2 * h)
    bad = (x.sum(axis=1) & 1) != p
    x[bad] = lex(2 * h[bad] + 1)
    return x


class Session:
    """QOP_init/QOP_finalize around one lattice, with numpy helpers."""

    def __init__(self, L):
        self.L = list(L)
        self.lib = lib()
        if not self.lib.QOPX_device_ok():
            raise RuntimeError("no HIP device: the MI355X library has no CPU path")
        ls = ivec(self.L)
        self._ls = ls
        lay = QOPLayout(None, None, 4, ls, 0, None, 0, volume(L))
        if self.lib.QOP_init(C.byref(lay)) != QOP_SUCCESS:
            raise RuntimeError("QOP_init failed: " + self.lib.QOPX_last_error().decode())
        self.V = volume(L)

    def close(self):
        if self.lib.QOP_is_initialized():
            self.lib.QOP_finalize()

    # vectors
    def vec(self, raw, eo=QOP_EVENODD, prec="D"):
        raw = np.ascontiguousarray(raw, dtype=np.float64 if prec == "D" else np.float32)
        f = self.lib.QOP_D3_create_V_from_raw if prec == "D" else self.lib.QOP_F3_create_V_from_raw
        return f(raw.reshape(-1), eo)

    def zeros(self, prec="D"):
        return self.vec(np.zeros((self.V, 3, 2)), QOP_EVENODD, prec)

    def get(self, v, eo=QOP_EVENODD, prec="D"):
        out = np.zeros((self.V, 3, 2), dtype=np.float64 if prec == "D" else np.float32)
        f = self.lib.QOP_D3_extract_V_to_raw if prec == "D" else self.lib.QOP_F3_extract_V_to_raw
        f(out.reshape(-1), v, eo)
        return out

    def free(self, v, prec="D"):
        (self.lib.QOP_D3_destroy_V if prec == "D" else self.lib.QOP_F3_destroy_V)(v)

    # links
    def links_from_raw(self, fat, lng, prec="D"):
        dt = np.float64 if prec == "D" else np.float32
        fat = [np.ascontiguousarray(f, dtype=dt) for f in fat]
        lng = [np.ascontiguousarray(l, dtype=dt) for l in lng] if lng is not None else None
        f = self.lib.QOP_D3_asqtad_create_L_from_raw if prec == "D" else self.lib.QOP_F3_asqtad_create_L_from_raw
        return f(ptr_array(fat), ptr_array(lng) if lng is not None else None, QOP_EVENODD)

    def gauge(self, U, prec="D"):
        dt = np.float64 if prec == "D" else np.float32
        U = [np.ascontiguousarray(u, dtype=dt) for u in U]
        f = self.lib.QOP_D3_create_G_from_raw if prec == "D" else self.lib.QOP_F3_create_G_from_raw
        return f(ptr_array(U), QOP_EVENODD)

    def force(self, F, prec="D"):
        """A QOP_Force from four raw (V, 3, 3, 2) arrays."""
        dt = np.float64 if prec == "D" else np.float32
        F = [np.ascontiguousarray(f, dtype=dt) for f in F]
        f = self.lib.QOP_D3_create_F_from_raw if prec == "D" else self.lib.QOP_F3_create_F_from_raw
        return f(ptr_array(F), QOP_EVENODD)

    def get_force(self, f, prec="D"):
        dt = np.float64 if prec == "D" else np.float32
        out = [np.zeros((self.V, 3, 3, 2), dtype=dt) for _ in range(4)]
        g = self.lib.QOP_D3_extract_F_to_raw if prec == "D" else self.lib.QOP_F3_extract_F_to_raw
        g(ptr_array(out), f, QOP_EVENODD)
        return out

    def asqtad_force(self, g, f, eps, xs, coeffs=None, prec="D"):
        """QOP_asqtad_force_multi(g, f, coeffs, eps, xs); returns the QOP_info_t."""
        c = QOPAsqtadCoeffs(**(coeffs or MILC_ASQTAD))
        info = QOPInfo()
        n = len(xs)
        ct = C.c_double if prec == "D" else C.c_float
        fn = self.lib.QOP_D3_asqtad_force_multi if prec == "D" else self.lib.QOP_F3_asqtad_force_multi
        fn(C.byref(info), g, f, C.byref(c), (ct * n)(*eps), (C.c_void_p * n)(*xs), n)
        return info

    def links_from_gauge(self, g, coeffs=None):
        c = QOPAsqtadCoeffs(**(coeffs or MILC_ASQTAD))
        info = QOPInfo()
        fla = self.lib.QOP_D3_asqtad_create_L_from_G(C.byref(info), C.byref(c), g)
        return fla, info

    def extract_links(self, fla, with_long=True):
        fat = [np.zeros((self.V, 3, 3, 2)) for _ in range(4)]
        lng = [np.zeros((self.V, 3, 3, 2)) for _ in range(4)] if with_long else None
        self.lib.QOP_D3_asqtad_extract_L_to_raw(ptr_array(fat), ptr_array(lng) if lng else None,
                                                fla, QOP_EVENODD)
        return fat, lng

    def set_opts(self, **kw):
        opts = (QOPOpt * len(kw))()
        keep = []
        for i, (k, v) in enumerate(kw.items()):
            b = k.encode()
            keep.append(b)
            opts[i] = QOPOpt(b, float(v), None)
        return self.lib.QOP_asqtad_invert_set_opts(opts, len(kw))


def milc_rephase_args(L, antiperiodic_t=True):
    """r0, bc and sign structs for MILC phases {0,1,3,7} and (1,1,1,-1) bc
    (examples/asqtad_test.c:275-281)."""
    r0 = ivec([0, 0, 0, 0])
    ph = (QOPComplex * 4)(QOPComplex(1, 0), QOPComplex(1, 0), QOPComplex(1, 0),
                          QOPComplex(-1 if antiperiodic_t else 1, 0))
    bc = QOPBc(ph)
    sm = ivec([0, 1, 3, 7])
    sign = QOPSign(sm)
    return r0, bc, sign, (ph, sm)


class Server:
    """QOPX_serve_* request server: resident links of several lattices and
    configurations, continuous batching of EVEN/ODD solves.  Source and
    output arrays are kept alive here until their request is waited for."""

    def __init__(self, capacity=16, prec=16, max_bytes=0):
        self.lib = lib()
        self.h = self.lib.QOPX_serve_create(capacity, prec, max_bytes)
        if not self.h:
            raise RuntimeError("QOPX_serve_create failed (no HIP device or bad arguments)")
        self.L = {}
        self._keep = {}

    def add_gauge(self, L, U, coeffs=None, milc_phases=True):
        U = [np.ascontiguousarray(u, dtype=np.float64) for u in U]
        c = QOPAsqtadCoeffs(**(coeffs or MILC_ASQTAD))
        if milc_phases:
            r0, bc, sign, keep = milc_rephase_args(L)
            cfg = self.lib.QOPX_serve_add_gauge(self.h, ivec(L), ptr_array(U), C.byref(c), r0,
                                                C.byref(bc), C.byref(sign))
        else:
            cfg = self.lib.QOPX_serve_add_gauge(self.h, ivec(L), ptr_array(U), C.byref(c), None,
                                                None, None)
        if cfg < 0:
            raise RuntimeError("configuration does not fit the HBM budget")
        self.L[cfg] = list(L)
        return cfg

    def add_links(self, L, fat, lng):
        fat = [np.ascontiguousarray(f, dtype=np.float64) for f in fat]
        lng = [np.ascontiguousarray(x, dtype=np.float64) for x in lng] if lng is not None else None
        cfg = self.lib.QOPX_serve_add_links(self.h, ivec(L), ptr_array(fat),
                                            ptr_array(lng) if lng is not None else None)
        if cfg < 0:
            raise RuntimeError("configuration does not fit the HBM budget")
        self.L[cfg] = list(L)
        return cfg

    def submit(self, cfg, src, mass, tol, eo=QOP_EVEN, max_iter=0, out=None):
        src = np.ascontiguousarray(src, dtype=np.float64)
        if out is None:
            out = np.zeros_like(src)
        rid = self.lib.QOPX_serve_submit(self.h, cfg, eo, mass, tol, max_iter,
                                         src.ctypes.data, out.ctypes.data)
        if rid < 0:
            raise RuntimeError("QOPX_serve_submit rejected the request")
        self._keep[rid] = (src, out)
        return rid

    def wait(self, rid):
        """-> (out, status, QOPResidArg, latency seconds)"""
        res = QOPResidArg()
        lat = C.c_double(0)
        st = self.lib.QOPX_serve_wait(self.h, rid, C.byref(res), C.byref(lat))
        src, out = self._keep.pop(rid)
        return out, st, res, lat.value

    def test(self, rid):
        return self.lib.QOPX_serve_test(self.h, rid)

    def stats(self):
        s = (C.c_double * 8)()
        self.lib.QOPX_serve_stats(self.h, s)
        keys = ("completed", "queued", "running", "batches", "bytes_used", "byte_budget",
                "chunks", "failed")
        return dict(zip(keys, list(s)))

    def close(self):
        if self.h:
            self.lib.QOPX_serve_destroy(self.h)
            self.h = None
