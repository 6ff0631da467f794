"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` (as the checker) and bench.py's
``cpu_baseline`` leg may import this module.  The compiled library is
``oracle/liboracle.so`` (built by ``oracle/Makefile``).
"""
import ctypes
import os

import numpy as np

from basis import System  # noqa: F401  (re-exported for callers)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class _Basis(ctypes.Structure):
    _fields_ = [("nshells", ctypes.c_int),
                ("am", ctypes.POINTER(ctypes.c_int)),
                ("nprim", ctypes.POINTER(ctypes.c_int)),
                ("prim_off", ctypes.POINTER(ctypes.c_int)),
                ("xyz", ctypes.POINTER(ctypes.c_double)),
                ("alpha", ctypes.POINTER(ctypes.c_double)),
                ("coef", ctypes.POINTER(ctypes.c_double))]


class _Shell(ctypes.Structure):
    _fields_ = [("am", ctypes.c_int), ("nprim", ctypes.c_int),
                ("xyz", ctypes.c_double * 3),
                ("alpha", ctypes.POINTER(ctypes.c_double)),
                ("coef", ctypes.POINTER(ctypes.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.ora_create.restype = vp
        L.ora_create.argtypes = [ctypes.POINTER(_Basis), ctypes.POINTER(_Basis), ctypes.c_int, dp, dp,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int]
        for fn in ("ora_setup", "ora_build_df_tensor", "ora_build_fock", "ora_build_density", "ora_diis"):
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = ctypes.c_int
        L.ora_d2cocc.argtypes = [vp]
        L.ora_d2cocc.restype = None
        L.ora_do_scf.argtypes = [vp, ctypes.c_int]
        L.ora_do_scf.restype = ctypes.c_int
        L.ora_calc_energy.argtypes = [vp]
        L.ora_calc_energy.restype = ctypes.c_double
        L.ora_destroy.argtypes = [vp]
        L.ora_destroy.restype = None
        L.ora_set_flags.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.ora_set_flags.restype = None
        L.ora_set_quirk_a2.argtypes = [vp, ctypes.c_int]
        L.ora_set_quirk_a2.restype = None
        L.ora_get_dptr.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
        L.ora_get_dptr.restype = dp
        L.ora_get_iptr.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
        L.ora_get_iptr.restype = ip
        L.ora_get_int.argtypes = [vp, ctypes.c_char_p]
        L.ora_get_int.restype = ctypes.c_int
        L.ora_set_int.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.ora_set_int.restype = None
        L.ora_get_double.argtypes = [vp, ctypes.c_char_p]
        L.ora_get_double.restype = ctypes.c_double
        L.ora_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.ora_boys.restype = None
        shp = ctypes.POINTER(_Shell)
        L.ora_overlap.argtypes = [shp, shp, dp]
        L.ora_kinetic.argtypes = [shp, shp, dp]
        L.ora_nuclear.argtypes = [shp, shp, ctypes.c_int, dp, dp, dp]
        L.ora_eri3.argtypes = [shp, shp, shp, dp]
        L.ora_eri2.argtypes = [shp, shp, dp]
        L.ora_eri4_elem.argtypes = [shp, shp, shp, shp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_eri4_elem.restype = ctypes.c_double
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_get_threads.restype = ctypes.c_int
        L.ora_lapack_is_mkl.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def boys(nmax, T):
    out = np.zeros(nmax + 1)
    lib().ora_boys(nmax, float(T), _dp(out))
    return out


class Shell:
    """A standalone contracted shell for integral unit tests."""

    def __init__(self, am, xyz, alpha, coef):
        self.alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64)
        self.s = _Shell(am, len(self.alpha), (ctypes.c_double * 3)(*xyz), _dp(self.alpha), _dp(self.coef))
        self.am = am

    @property
    def n(self):
        return (self.am + 1) * (self.am + 2) // 2


def overlap(A, B):
    out = np.zeros(A.n * B.n)
    lib().ora_overlap(ctypes.byref(A.s), ctypes.byref(B.s), _dp(out))
    return out.reshape(A.n, B.n)


def kinetic(A, B):
    out = np.zeros(A.n * B.n)
    lib().ora_kinetic(ctypes.byref(A.s), ctypes.byref(B.s), _dp(out))
    return out.reshape(A.n, B.n)


def nuclear(A, B, Z, xyz):
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    out = np.zeros(A.n * B.n)
    lib().ora_nuclear(ctypes.byref(A.s), ctypes.byref(B.s), len(Z), _dp(Z), _dp(xyz), _dp(out))
    return out.reshape(A.n, B.n)


def eri3(A, B, C):
    out = np.zeros(A.n * B.n * C.n)
    lib().ora_eri3(ctypes.byref(A.s), ctypes.byref(B.s), ctypes.byref(C.s), _dp(out))
    return out.reshape(A.n, B.n, C.n)


def eri2(C, D):
    out = np.zeros(C.n * D.n)
    lib().ora_eri2(ctypes.byref(C.s), ctypes.byref(D.s), _dp(out))
    return out.reshape(C.n, D.n)


def eri4(A, B, C, D, ia, ib, ic, id_):
    return lib().ora_eri4_elem(ctypes.byref(A.s), ctypes.byref(B.s), ctypes.byref(C.s),
                               ctypes.byref(D.s), ia, ib, ic, id_)


def set_threads(n):
    lib().ora_set_threads(int(n))


def get_threads():
    return lib().ora_get_threads()


def is_mkl():
    return bool(lib().ora_lapack_is_mkl())


class OracleSCF:
    """The reference SCF engine restated on the CPU, driven from Python."""

    def __init__(self, system, niters=50, quirk_a5=True, quirk_a1=True, ene_tol=1e-11, scrtol=1e-11,
                 quirk_a2=True):
        L = lib()
        self.sys = system
        self._keep = []
        bas = self._mk_basis(system.basis)
        aux = self._mk_basis(system.auxbasis)
        Z = np.ascontiguousarray(system.mol["Z"])
        xyz = np.ascontiguousarray(system.mol["coords"].reshape(-1))
        sad = np.ascontiguousarray(system.sad)
        self._keep += [Z, xyz, sad]
        self.h = L.ora_create(ctypes.byref(bas), ctypes.byref(aux), system.mol["natoms"], _dp(Z), _dp(xyz),
                              system.mol["charge"], system.mol["nelectrons"], system.mol["enuc"],
                              _dp(sad), niters)
        L.ora_set_flags(self.h, int(quirk_a5), int(quirk_a1), ene_tol, scrtol)
        L.ora_set_quirk_a2(self.h, int(quirk_a2))

    def _mk_basis(self, b):
        arrs = [np.ascontiguousarray(b[k]) for k in ("am", "nprim", "prim_off")]
        xyz = np.ascontiguousarray(b["xyz"].reshape(-1))
        al = np.ascontiguousarray(b["alpha"])
        co = np.ascontiguousarray(b["coef"])
        self._keep += arrs + [xyz, al, co]
        s = _Basis(b["nshells"], _ip(arrs[0]), _ip(arrs[1]), _ip(arrs[2]), _dp(xyz), _dp(al), _dp(co))
        self._keep.append(s)
        return s

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ora_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def setup(self):
        return lib().ora_setup(self.h)

    def build_df_tensor(self):
        return lib().ora_build_df_tensor(self.h)

    def build_fock(self):
        return lib().ora_build_fock(self.h)

    def build_density(self):
        return lib().ora_build_density(self.h)

    def diis(self):
        return lib().ora_diis(self.h)

    def d2cocc(self):
        lib().ora_d2cocc(self.h)

    def calc_energy(self):
        return lib().ora_calc_energy(self.h)

    def do_scf(self, verbose=False):
        return lib().ora_do_scf(self.h, int(verbose))

    def d(self, name, shape=None):
        n = ctypes.c_size_t(0)
        p = lib().ora_get_dptr(self.h, name.encode(), ctypes.byref(n))
        if not p or n.value == 0:
            return np.zeros(0)
        a = np.ctypeslib.as_array(p, shape=(n.value,))
        return a.reshape(shape) if shape is not None else a

    def i(self, name):
        n = ctypes.c_size_t(0)
        p = lib().ora_get_iptr(self.h, name.encode(), ctypes.byref(n))
        if not p or n.value == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def geti(self, name):
        return lib().ora_get_int(self.h, name.encode())

    def seti(self, name, v):
        lib().ora_set_int(self.h, name.encode(), int(v))

    def getd(self, name):
        return lib().ora_get_double(self.h, name.encode())

    def mat(self, name):
        nbf = self.geti("nbf")
        return self.d(name, (nbf, nbf))
