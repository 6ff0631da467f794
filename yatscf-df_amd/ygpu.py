"""ctypes bindings of the MI355X DF-RHF path.

* :class:`Engine` -- the C-ABI of ``libyatscf_gpu.so`` (include/yatscf_gpu.h),
  one device engine per instance.
* :class:`TinySCF` -- the TinySCF host engine ``libtinyscf.so`` (the drop-in
  for the reference's TinySCF API) driven step by step.

Both raise :class:`GpuError` on any non-zero status.  There is no fallback:
if the shared libraries are missing, importing them fails loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(HERE, "libyatscf_gpu.so")
HOST_LIB = os.path.join(HERE, "libtinyscf.so")
DATA = os.path.join(HERE, "data")
BASIS = os.path.join(DATA, "cc-pvdz", "cc-pvdz.gbs")
AUXBASIS = os.path.join(DATA, "cc-pVDZ-RI", "cc-pVDZ-RI.gbs")

YGPU_OK = 0
ERRORS = {1: "INVALID", 2: "HIP", 3: "NOMEM", 4: "SOLVER", 5: "NUMERIC", 6: "NODEVICE"}

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)
vp = ctypes.c_void_p


class GpuError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(status, "?"), status, msg))
        self.status = status


class DfTiming(ctypes.Structure):
    _fields_ = [("t_3c", ctypes.c_double), ("t_2c", ctypes.c_double),
                ("t_invsqrt", ctypes.c_double), ("t_dfgemm", ctypes.c_double)]


class FockTiming(ctypes.Structure):
    _fields_ = [("t_tempJ", ctypes.c_double), ("t_Jmat", ctypes.c_double), ("t_tempK", ctypes.c_double),
                ("t_Kmat", ctypes.c_double), ("t_symm", ctypes.c_double), ("t_total", ctypes.c_double)]


# name -> argtypes (all return int status)
SIGNATURES = {
    "ygpu_device_count": [ip],
    "ygpu_create": [ctypes.c_int, ctypes.POINTER(vp)],
    "ygpu_destroy": [vp],
    "ygpu_release_cached_memory": [],
    "ygpu_set_basis": [vp, ctypes.c_int, ctypes.c_int, ip, ip, ip, dp, dp, dp],
    "ygpu_set_atoms": [vp, ctypes.c_int, dp, dp],
    "ygpu_ovlp_hcore": [vp, dp, dp],
    "ygpu_sym_inv_sqrt": [vp, ctypes.c_int, dp, dp, dp],
    "ygpu_schwarz": [vp, ctypes.c_int, dp, dp, dp],
    "ygpu_set_sparsity": [vp, ip, ctypes.c_int, ip, ip, dp, dp, ctypes.c_double],
    "ygpu_build_df_tensor": [vp, dp, dp, ctypes.POINTER(DfTiming)],
    "ygpu_build_df_tensor_ex": [vp, dp, dp, dp, ctypes.POINTER(DfTiming)],
    "ygpu_set_df_tensor": [vp, dp],
    "ygpu_build_fock": [vp, ctypes.c_int, dp, dp, dp, dp, dp, dp, ctypes.POINTER(FockTiming)],
    "ygpu_build_fock_noalloc": [vp, ctypes.c_int, dp, dp, dp, dp, dp, dp],
    "ygpu_build_density": [vp, ctypes.c_int, dp, dp, dp, dp, dp, ip],
    "ygpu_dgemm": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int],
    "ygpu_layout_info": [vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long), ip],
    "ygpu_set_overlap": [vp, dp, dp],
    "ygpu_diis_transform": [vp, dp, dp],
    "ygpu_diis_update": [vp, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp],
    "ygpu_diis_extrapolate": [vp, ctypes.c_int, dp, dp],
    "ygpu_diis_error": [vp, ctypes.c_int, dp],
    "ygpu_set_fock_precision": [vp, ctypes.c_int],
    "ygpu_set_metric_factor": [vp, ctypes.c_int],
    "ygpu_sph_coefficients": [ctypes.c_int, dp],
    "ygpu_set_aux_spherical": [vp, ctypes.c_int],
    "ygpu_set_orth_dim": [vp, ctypes.c_int],
    "ygpu_fock_last_info": [vp, ip, dp],
    "ygpu_set_density_method": [vp, ctypes.c_int],
    "ygpu_density_last_info": [vp, ip, ip],
    "ygpu_mlp_forward": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "ygpu_mlp_trainer_create": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(vp)],
    "ygpu_mlp_trainer_epoch": [vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint, dp],
    "ygpu_mlp_trainer_grad": [vp, ctypes.c_int, ip, ctypes.POINTER(ctypes.c_float), dp],
    "ygpu_mlp_trainer_params": [vp, ctypes.POINTER(ctypes.c_float)],
    "ygpu_mlp_trainer_destroy": [vp],
}

FOCK_FP64, FOCK_MIXED, FOCK_MIXED6, FOCK_INCR = 0, 1, 2, 3
DENSITY_EIGEN, DENSITY_PURIFY = 0, 1
METRIC_CHOLESKY, METRIC_INVSQRT = 0, 1

fp = ctypes.POINTER(ctypes.c_float)


class YsmlMol(ctypes.Structure):
    """ysml_mol_t of host/guess_ml.h: the molecule as the pair model sees it."""
    _fields_ = [("natoms", ctypes.c_int), ("nbf", ctypes.c_int), ("Z", dp), ("xyz", dp),
                ("atom_shell_start", ip), ("am", ip), ("shell_bf_start", ip)]


# the pair model's dimensions (host/guess_ml.h)
YSML_NCEN, YSML_NEL, YSML_NTYPES, YSML_NOUT = 32, 2, 5, 29
YSML_NIN = 2 * YSML_NEL * YSML_NCEN + YSML_NCEN + YSML_NTYPES

_glib = None
_hlib = None


def gpu_lib():
    global _glib
    if _glib is None:
        if not os.path.exists(GPU_LIB):
            raise RuntimeError("libyatscf_gpu.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(GPU_LIB, mode=ctypes.RTLD_GLOBAL)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.ygpu_last_error.restype = ctypes.c_char_p
        L.ygpu_last_error.argtypes = []
        _glib = L
    return _glib


def host_lib():
    global _hlib
    if _hlib is None:
        gpu_lib()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("libtinyscf.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(HOST_LIB)
        L.tinyscf_alloc.restype = vp
        L.tinyscf_dptr.restype = dp
        L.tinyscf_dptr.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
        L.tinyscf_iptr.restype = ip
        L.tinyscf_iptr.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
        L.tinyscf_geti.argtypes = [vp, ctypes.c_char_p]
        L.tinyscf_geti.restype = ctypes.c_int
        L.tinyscf_seti.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.tinyscf_seti.restype = None
        L.tinyscf_getd.argtypes = [vp, ctypes.c_char_p]
        L.tinyscf_getd.restype = ctypes.c_double
        L.tinyscf_setd.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
        L.tinyscf_setd.restype = None
        L.tinyscf_call.argtypes = [vp, ctypes.c_char_p]
        L.tinyscf_call.restype = ctypes.c_int
        L.tinyscf_error.argtypes = []
        L.tinyscf_error.restype = ctypes.c_char_p
        L.tinyscf_gpu.argtypes = [vp]
        L.tinyscf_fill_mol.argtypes = [vp, ctypes.POINTER(YsmlMol), ip]
        L.tinyscf_fill_mol.restype = None
        L.tinyscf_set_guess_weights.argtypes = [vp, ctypes.c_char_p]
        L.tinyscf_set_guess_weights.restype = None
        L.ysml_supported.argtypes = [ctypes.POINTER(YsmlMol)]
        L.ysml_pairs.argtypes = [ctypes.POINTER(YsmlMol), ip, ctypes.c_int]
        L.ysml_features.argtypes = [ctypes.POINTER(YsmlMol), ctypes.c_int, ip, fp]
        L.ysml_targets.argtypes = [ctypes.POINTER(YsmlMol), ctypes.c_int, ip, dp, fp, fp]
        L.ysml_assemble.argtypes = [ctypes.POINTER(YsmlMol), ctypes.c_int, ip, fp, dp]
        for fn in ("ysml_supported", "ysml_pairs", "ysml_features", "ysml_targets", "ysml_assemble"):
            getattr(L, fn).restype = ctypes.c_int
        L.tinyscf_gpu.restype = vp
        L.TinySCF_try_init.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.TinySCF_try_init.restype = ctypes.c_int
        for fn in ("TinySCF_init_batch_dgemm_arrays", "TinySCF_compute_Hcore_Ovlp_mat",
                   "TinySCF_compute_sq_Schwarz_scrvals", "TinySCF_prepare_sparsity", "TinySCF_get_initial_guess",
                   "TinySCF_build_DF_tensor", "TinySCF_do_SCF", "TinySCF_build_FockMat", "TinySCF_build_DenMat",
                   "TinySCF_DIIS", "TinySCF_D2Cocc", "TinySCF_free_batch_dgemm_arrays", "free_TinySCF"):
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = None
        L.quickSort_eigval.argtypes = [dp, ip, ctypes.c_int, ctypes.c_int]
        L.quickSort_eigval.restype = None
        _hlib = L
    return _hlib


def last_error():
    return gpu_lib().ygpu_last_error().decode()


def check(status):
    if status != YGPU_OK:
        raise GpuError(status, last_error())


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(ip)


def release_cached_memory():
    """Return the library's cache of freed device blocks to the driver."""
    check(gpu_lib().ygpu_release_cached_memory())


def device_count():
    n = ctypes.c_int(0)
    gpu_lib().ygpu_device_count(ctypes.byref(n))
    return n.value


class Engine:
    """One device engine (the C-ABI handle)."""

    def __init__(self, device=0):
        self.L = gpu_lib()
        h = vp()
        check(self.L.ygpu_create(device, ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            st = self.L.ygpu_destroy(self.h)
            self.h = None
            check(st)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_basis(self, which, b):
        arrs = [np.ascontiguousarray(b[k], dtype=np.int32) for k in ("am", "nprim", "prim_off")]
        fl = [np.ascontiguousarray(b[k], dtype=np.float64).reshape(-1) for k in ("xyz", "alpha", "coef")]
        self._keep["basis%d" % which] = arrs + fl
        check(self.L.ygpu_set_basis(self.h, which, int(b["nshells"]), _i(arrs[0]), _i(arrs[1]), _i(arrs[2]),
                                    _d(fl[0]), _d(fl[1]), _d(fl[2])))

    def set_atoms(self, Z, xyz):
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        self._keep["atoms"] = (Z, xyz)
        check(self.L.ygpu_set_atoms(self.h, len(Z), _d(Z), _d(xyz)))

    def set_system(self, system):
        """Upload an oracle.basis.System (both shell sets and the nuclei)."""
        self.nbf, self.naux = system.nbf, system.naux
        self.nshells, self.aux_nshells = system.basis["nshells"], system.auxbasis["nshells"]
        self.set_basis(0, system.basis)
        self.set_basis(1, system.auxbasis)
        self.set_atoms(system.mol["Z"], system.mol["coords"])

    def ovlp_hcore(self):
        S = np.zeros((self.nbf, self.nbf))
        H = np.zeros((self.nbf, self.nbf))
        check(self.L.ygpu_ovlp_hcore(self.h, _d(S), _d(H)))
        return S, H

    def sym_inv_sqrt(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        X = np.zeros_like(A)
        m = ctypes.c_double(0)
        check(self.L.ygpu_sym_inv_sqrt(self.h, A.shape[0], _d(A), _d(X), ctypes.byref(m)))
        return X, m.value

    def schwarz(self, quirk_a5=True):
        sp = np.zeros(self.nshells * self.nshells)
        bf = np.zeros(self.nbf * self.nbf)
        df = np.zeros(self.aux_nshells)
        check(self.L.ygpu_schwarz(self.h, int(quirk_a5), _d(sp), _d(bf), _d(df)))
        return sp, bf, df

    def set_sparsity(self, mask, lid, rid, sp, df, scrtol2=1e-22):
        mask = np.ascontiguousarray(mask, dtype=np.int32).reshape(-1)
        lid = np.ascontiguousarray(lid, dtype=np.int32)
        rid = np.ascontiguousarray(rid, dtype=np.int32)
        sp = np.ascontiguousarray(sp, dtype=np.float64)
        df = np.ascontiguousarray(df, dtype=np.float64)
        self._keep["sparsity"] = (mask, lid, rid, sp, df)
        self.nnz = int(np.sum(mask >= 0))
        check(self.L.ygpu_set_sparsity(self.h, _i(mask), len(lid), _i(lid), _i(rid), _d(sp), _d(df), scrtol2))

    def build_df_tensor(self, want_pqA=False, want_Jpq=False, want_B=False):
        pqA = np.zeros((self.nnz, self.naux)) if want_pqA else None
        Jpq = np.zeros((self.naux, self.naux)) if want_Jpq else None
        B = np.zeros((self.nnz, self.naux)) if want_B else None
        t = DfTiming()
        check(self.L.ygpu_build_df_tensor_ex(self.h, _d(pqA), _d(Jpq), _d(B), ctypes.byref(t)))
        return dict(pqA=pqA, Jpq=Jpq, B=B, timing=t)

    def fock_last_info(self):
        """(arithmetic the last Fock build ran, max |Q^T Q - I| of a difference build's alignment)"""
        m, e = ctypes.c_int(0), ctypes.c_double(0.0)
        check(self.L.ygpu_fock_last_info(self.h, ctypes.byref(m), ctypes.byref(e)))
        return m.value, e.value

    def set_metric_factor(self, mode):
        """METRIC_CHOLESKY (W = L^-1, the default) or METRIC_INVSQRT (W = J^-1/2, the reference's)."""
        check(self.L.ygpu_set_metric_factor(self.h, int(mode)))

    def set_df_tensor(self, B):
        B = np.ascontiguousarray(B, dtype=np.float64)
        check(self.L.ygpu_set_df_tensor(self.h, _d(B)))

    def build_fock(self, D, Cocc, H, want_JK=True, n_occ=None):
        """D = Cocc = None: the last density build's D and C_occ, already on the device (give n_occ);
        H = None: the copy cached by ovlp_hcore."""
        if Cocc is not None:
            n_occ = Cocc.shape[1]
            D = np.ascontiguousarray(D, dtype=np.float64)
            Cocc = np.ascontiguousarray(Cocc, dtype=np.float64)
        if H is not None:
            H = np.ascontiguousarray(H, dtype=np.float64)
        F = np.zeros((self.nbf, self.nbf))
        J = np.zeros((self.nbf, self.nbf)) if want_JK else None
        K = np.zeros((self.nbf, self.nbf)) if want_JK else None
        t = FockTiming()
        check(self.L.ygpu_build_fock(self.h, n_occ, _d(D), _d(Cocc), _d(H), _d(J), _d(K), _d(F), ctypes.byref(t)))
        return dict(F=F, J=J, K=K, timing=t)

    def build_fock_noalloc(self, D, Cocc, H, want_JK=True):
        """ygpu_build_fock_noalloc: the FP64 Fock build with its device buffers in the library's
        static scratch array (no device allocation); host D, C_occ and Hcore are required."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        Cocc = np.ascontiguousarray(Cocc, dtype=np.float64)
        H = np.ascontiguousarray(H, dtype=np.float64)
        F = np.zeros((self.nbf, self.nbf))
        J = np.zeros((self.nbf, self.nbf)) if want_JK else None
        K = np.zeros((self.nbf, self.nbf)) if want_JK else None
        check(self.L.ygpu_build_fock_noalloc(self.h, Cocc.shape[1], _d(D), _d(Cocc), _d(H), _d(J), _d(K), _d(F)))
        return dict(F=F, J=J, K=K)

    def set_fock_precision(self, mode):
        """FOCK_FP64 (reference arithmetic) or FOCK_MIXED (BF16-split K, incremental FP32-tensor J)."""
        check(self.L.ygpu_set_fock_precision(self.h, int(mode)))

    def build_density(self, Fo, X, n_occ):
        """Fo = None: the last DIIS transform / extrapolation on the device; X = None: the cached X."""
        if Fo is not None:
            Fo = np.ascontiguousarray(Fo, dtype=np.float64)
        if X is not None:
            X = np.ascontiguousarray(X, dtype=np.float64)
        n = Fo.shape[0] if Fo is not None else self.nbf
        D = np.zeros((n, n))
        C = np.zeros((n, n_occ))
        ev = np.zeros(n)
        idx = np.zeros(n, dtype=np.int32)
        check(self.L.ygpu_build_density(self.h, n_occ, _d(Fo), _d(X), _d(D), _d(C), _d(ev), _i(idx)))
        return D, C, ev, idx

    def set_density_method(self, method):
        """DENSITY_EIGEN (default) or DENSITY_PURIFY for the following build_density calls."""
        check(self.L.ygpu_set_density_method(self.h, int(method)))

    def density_last_info(self):
        """(method used, SP2 steps) of the last build_density."""
        m, k = ctypes.c_int(), ctypes.c_int()
        check(self.L.ygpu_density_last_info(self.h, ctypes.byref(m), ctypes.byref(k)))
        return m.value, k.value

    def dgemm(self, ta, tb, A, B, alpha=1.0, beta=0.0, C=None):
        A = np.ascontiguousarray(A, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        m = A.shape[1] if ta else A.shape[0]
        k = A.shape[0] if ta else A.shape[1]
        n = B.shape[0] if tb else B.shape[1]
        C = np.zeros((m, n)) if C is None else np.ascontiguousarray(C, dtype=np.float64).copy()
        check(self.L.ygpu_dgemm(self.h, int(ta), int(tb), m, n, k, alpha, _d(A), A.shape[1], _d(B), B.shape[1],
                                beta, _d(C), n))
        return C

    def set_overlap(self, S, X):
        S = np.ascontiguousarray(S, dtype=np.float64)
        X = np.ascontiguousarray(X, dtype=np.float64)
        self._keep["overlap"] = (S, X)
        check(self.L.ygpu_set_overlap(self.h, _d(S), _d(X)))

    def diis_transform(self, F):
        """F = None: the last Fock build's F on the device."""
        if F is not None:
            F = np.ascontiguousarray(F, dtype=np.float64)
        n = F.shape[0] if F is not None else self.nbf
        out = np.zeros((n, n))
        check(self.L.ygpu_diis_transform(self.h, _d(F), _d(out)))
        return out

    def diis_update(self, slot, length, F, D, want_forth=True):
        """F = D = None: the last Fock build's F and its density on the device."""
        if F is not None:
            F = np.ascontiguousarray(F, dtype=np.float64)
        if D is not None:
            D = np.ascontiguousarray(D, dtype=np.float64)
        dots = np.zeros(length)
        n = F.shape[0] if F is not None else self.nbf
        out = np.zeros((n, n)) if want_forth else None
        check(self.L.ygpu_diis_update(self.h, slot, length, _d(F), _d(D), _d(dots), _d(out)))
        return dots, out

    def diis_error(self, slot):
        v = ctypes.c_double(0.0)
        check(self.L.ygpu_diis_error(self.h, slot, ctypes.byref(v)))
        return v.value

    def diis_extrapolate(self, coef):
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        out = np.zeros((self.nbf, self.nbf))
        check(self.L.ygpu_diis_extrapolate(self.h, len(coef), _d(coef), _d(out)))
        return out

    def layout_info(self):
        a, b, c = ctypes.c_long(0), ctypes.c_long(0), ctypes.c_int(0)
        check(self.L.ygpu_layout_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value


class TinySCF:
    """The TinySCF host engine (libtinyscf.so) with the reference's call sequence."""

    def __init__(self, xyz, niters=50, basis=BASIS, auxbasis=AUXBASIS, verbose=False, basis_type=None):
        """basis_type: None (the .gbs files' own flags), "cartesian" or "spherical" (both basis
        sets)."""
        self.L = host_lib()
        self.T = self.L.tinyscf_alloc()
        os.environ.setdefault("YATSCF_VERBOSE", "1" if verbose else "0")
        if basis_type not in (None, "cartesian", "spherical"):
            raise ValueError("basis_type: None, 'cartesian' or 'spherical'")
        old = os.environ.get("YATSCF_BASIS_TYPE")
        if basis_type:
            os.environ["YATSCF_BASIS_TYPE"] = basis_type
        else:
            os.environ.pop("YATSCF_BASIS_TYPE", None)
        try:
            st = self.L.TinySCF_try_init(self.T, basis.encode(), auxbasis.encode(), xyz.encode(), niters)
        finally:
            if old is None:
                os.environ.pop("YATSCF_BASIS_TYPE", None)
            else:
                os.environ["YATSCF_BASIS_TYPE"] = old
        if st != 0:
            raise RuntimeError("TinySCF init failed (%d): %s" % (st, last_error()))
        self.L.tinyscf_seti(self.T, b"verbose", int(verbose))

    def step(self, name):
        st = self.L.tinyscf_call(self.T, name.encode())
        if st != 0:
            raise GpuError(st, "%s: %s" % (name, self.L.tinyscf_error().decode()))

    def setup(self):
        for fn in ("TinySCF_init_batch_dgemm_arrays", "TinySCF_compute_Hcore_Ovlp_mat",
                   "TinySCF_compute_sq_Schwarz_scrvals", "TinySCF_prepare_sparsity", "TinySCF_get_initial_guess"):
            self.step(fn)

    def build_df_tensor(self):
        self.step("TinySCF_build_DF_tensor")

    def do_scf(self):
        self.step("TinySCF_do_SCF")

    def build_fock(self, mode=FOCK_FP64):
        """One Fock build from the engine's current D and C_occ in the given arithmetic
        (FOCK_FP64 or FOCK_MIXED); the engine is left in FP64 mode.  Returns F (AO basis)."""
        g = self.L.tinyscf_gpu(self.T)
        L = gpu_lib()
        check(L.ygpu_set_fock_precision(g, int(mode)))
        try:
            self.step("TinySCF_build_FockMat")
        finally:
            check(L.ygpu_set_fock_precision(g, FOCK_FP64))
        return self.mat("F_mat").copy()

    def close(self):
        if getattr(self, "T", None):
            self.L.TinySCF_free_batch_dgemm_arrays(self.T)
            self.L.free_TinySCF(self.T)
            self.T = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def d(self, name, shape=None):
        n = ctypes.c_size_t(0)
        p = self.L.tinyscf_dptr(self.T, name.encode(), ctypes.byref(n))
        if not p or n.value == 0:
            return np.zeros(0)
        a = np.ctypeslib.as_array(p, shape=(n.value,))
        return a.reshape(shape) if shape is not None else a

    def i(self, name):
        n = ctypes.c_size_t(0)
        p = self.L.tinyscf_iptr(self.T, name.encode(), ctypes.byref(n))
        if not p or n.value == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def geti(self, name):
        return self.L.tinyscf_geti(self.T, name.encode())

    def seti(self, name, v):
        self.L.tinyscf_seti(self.T, name.encode(), int(v))

    def getd(self, name):
        return self.L.tinyscf_getd(self.T, name.encode())

    def setd(self, name, v):
        self.L.tinyscf_setd(self.T, name.encode(), float(v))

    def mat(self, name):
        n = self.geti("nbasfuncs")
        return self.d(name, (n, n))


def solve(xyz, niters=80, ene_tol=1e-11, quirk_a2=1, mixed=False, weights=None, fock_check=False,
          final_fock=False, schedule=None, keep_diis=None, reref=None, confirm=None, purify=None,
          verbose=False, basis_type=None,
          basis=BASIS, auxbasis=AUXBASIS):
    """One complete SCF of ``xyz`` through the TinySCF engine, timed from after input parsing
    (basis shells and nuclei uploaded) to the passing convergence test: S/H/X, Schwarz
    screening, pair layout, initial guess (SAD, plus the learned pair correction when
    ``weights`` names a network file), DF tensor, SCF iterations.  ``mixed`` starts the SCF in
    the reduced-precision Fock mode.  After the timer stops (untimed): with ``final_fock`` the
    FP64 Fock matrix of the converged density (AO basis) as "F" (the engine's F_mat holds the
    DIIS output in the orthogonal basis); with ``fock_check`` also one mixed Fock build on the
    same density, the reduced-precision arithmetic's own error in F and in the orbital energies
    of X^T F X.  ``basis_type``: see TinySCF.  Returns a dict ("nbf" the Cartesian components
    the matrices are written in, "nbf_sph" the orbital basis functions of the calculation); the
    engine is closed."""
    import time
    T = TinySCF(xyz, niters=niters, verbose=verbose, basis_type=basis_type, basis=basis, auxbasis=auxbasis)
    try:
        T.L.tinyscf_set_guess_weights(T.T, (weights or "").encode())
        T.setd("ene_tol", ene_tol)
        T.seti("quirk_a2", int(quirk_a2))
        T.seti("mixed", int(mixed))
        if schedule is not None:
            T.seti("mixed_schedule", int(schedule))
        if keep_diis is not None:
            T.seti("mixed_keep_diis", int(keep_diis))
        if reref is not None:
            T.setd("mixed_reref", float(reref))
        if confirm is not None:
            T.seti("mixed_confirm", int(confirm))
        if purify is not None:
            T.seti("purify", int(purify))
        t0 = time.perf_counter()
        T.setup()
        t1 = time.perf_counter()
        T.build_df_tensor()
        t2 = time.perf_counter()
        T.do_scf()
        t3 = time.perf_counter()
        nbf = T.geti("nbasfuncs")
        res = dict(seconds=t3 - t0, setup_s=t1 - t0, df_s=t2 - t1, scf_s=t3 - t2, iters=T.geti("iter"),
                   energy=T.getd("total_energy"), converged=bool(T.geti("converged")),
                   hist=T.d("t_hist").reshape(-1, 6).copy(), t_df=T.d("t_df").copy(), nbf=nbf,
                   naux=T.geti("df_nbf"), n_occ=T.geti("n_occ"), nnz=T.geti("num_bf_pair_scr"),
                   switch=T.geti("mixed_switch_iter"), switch6=T.geti("mixed6_switch_iter"),
                   phases=T.i("phase_log").copy(),
                   guess_ml=T.geti("guess_ml"), nbf_sph=T.geti("nbf_sph"), sph_orb=T.geti("sph_orb"),
                   sph_aux=T.geti("sph_aux"),
                   D=T.mat("D_mat").copy(), S=T.mat("S_mat").copy(),
                   X=T.mat("X_mat").copy(), H=T.mat("Hcore_mat").copy(), eigval=T.d("eigval").copy())
        mask = T.i("bf_pair_mask").reshape(nbf, nbf)
        res["n_up"] = int(np.sum(mask[np.triu_indices(nbf)] >= 0))
        if final_fock or fock_check:
            res["F"] = T.build_fock(FOCK_FP64)
        if fock_check:
            X, F64 = res["X"], res["F"]
            Fmx = T.build_fock(FOCK_MIXED)
            m = res["nbf_sph"]                  # X's columns beyond the orbital basis are zero
            e64 = np.linalg.eigvalsh((X.T @ F64 @ X)[:m, :m])
            emx = np.linalg.eigvalsh((X.T @ Fmx @ X)[:m, :m])
            no = res["n_occ"]
            res["fock_check"] = {"max_abs_dF": float(np.max(np.abs(Fmx - F64))),
                                 "max_abs_d_orbital_energy": float(np.max(np.abs(emx - e64))),
                                 "max_abs_d_occupied_orbital_energy": float(np.max(np.abs(emx[:no] - e64[:no])))}
        return res
    finally:
        T.close()


# ---------------------------------------------------------------------------
# learned-guess pair model (host/guess_ml.c) and its network (C-ABI)

def _f(a):
    return a.ctypes.data_as(fp) if a is not None else None


class PairModel:
    """Features, targets and the rebuild of the pair model for one molecule.

    Built from plain arrays (Z, xyz in Bohr, per-atom shell starts, shell angular
    momenta, shell first basis functions) or from a TinySCF engine."""

    def __init__(self, Z, xyz, atom_shell_start, am, shell_bf_start, nbf):
        self._keep = [np.ascontiguousarray(Z, dtype=np.float64), np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1),
                      np.ascontiguousarray(atom_shell_start, dtype=np.int32), np.ascontiguousarray(am, dtype=np.int32),
                      np.ascontiguousarray(shell_bf_start, dtype=np.int32)]
        z, x, ss, a, sb = self._keep
        self.mol = YsmlMol(len(z), int(nbf), _d(z), _d(x), _i(ss), _i(a), _i(sb))
        self.L = host_lib()
        n = self.L.ysml_pairs(ctypes.byref(self.mol), None, 0)
        self.pairs = np.zeros((n, 2), dtype=np.int32)
        self.L.ysml_pairs(ctypes.byref(self.mol), _i(self.pairs), n)

    @classmethod
    def of_engine(cls, T):
        m = YsmlMol()
        ss = np.zeros(T.geti("natoms") + 1, dtype=np.int32)
        T.L.tinyscf_fill_mol(T.T, ctypes.byref(m), _i(ss))
        na, ns = m.natoms, int(ss[-1])
        Z = np.ctypeslib.as_array(m.Z, shape=(na,)).copy()
        xyz = np.ctypeslib.as_array(m.xyz, shape=(3 * na,)).copy()
        am = np.ctypeslib.as_array(m.am, shape=(ns,)).copy()
        sb = np.ctypeslib.as_array(m.shell_bf_start, shape=(ns,)).copy()
        return cls(Z, xyz, ss, am, sb, m.nbf)

    def supported(self):
        return bool(self.L.ysml_supported(ctypes.byref(self.mol)))

    def types(self):
        X = self.features()
        return np.argmax(X[:, YSML_NIN - YSML_NTYPES:], axis=1)

    def features(self):
        X = np.zeros((len(self.pairs), YSML_NIN), dtype=np.float32)
        if self.L.ysml_features(ctypes.byref(self.mol), len(self.pairs), _i(self.pairs), _f(X)) != 0:
            raise RuntimeError("ysml_features failed")
        return X

    def targets(self, dD):
        dD = np.ascontiguousarray(dD, dtype=np.float64)
        Y = np.zeros((len(self.pairs), YSML_NOUT), dtype=np.float32)
        M = np.zeros_like(Y)
        if self.L.ysml_targets(ctypes.byref(self.mol), len(self.pairs), _i(self.pairs), _d(dD), _f(Y), _f(M)) != 0:
            raise RuntimeError("ysml_targets failed (unsupported shells)")
        return Y, M

    def assemble(self, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        n = self.mol.nbf
        dD = np.zeros((n, n))
        if self.L.ysml_assemble(ctypes.byref(self.mol), len(self.pairs), _i(self.pairs), _f(Y), _d(dD)) != 0:
            raise RuntimeError("ysml_assemble failed (unsupported shells)")
        return dD


def mlp_param_count(n_in, nh, n_out):
    return n_in * nh + nh + 2 * (nh * nh + nh) + nh * n_out + n_out


def mlp_forward(engine, n_in, nh, n_out, params, X):
    """Y = network(X) on the device of ``engine`` (an Engine)."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((X.shape[0], n_out), dtype=np.float32)
    check(engine.L.ygpu_mlp_forward(engine.h, n_in, nh, n_out, _f(params), X.shape[0], _f(X), _f(Y)))
    return Y


class MLPTrainer:
    """Device-resident training of the pair network (ygpu_mlp_trainer_*)."""

    def __init__(self, engine, nh, X, Y, M, params):
        self.engine = engine
        self.L = engine.L
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        M = np.ascontiguousarray(M, dtype=np.float32)
        params = np.ascontiguousarray(params, dtype=np.float32)
        self.n_in, self.nh, self.n_out, self.n = X.shape[1], nh, Y.shape[1], X.shape[0]
        self.np = mlp_param_count(self.n_in, nh, self.n_out)
        assert params.size == self.np and M.shape == Y.shape and Y.shape[0] == self.n
        self.h = vp()
        check(self.L.ygpu_mlp_trainer_create(engine.h, self.n_in, nh, self.n_out, self.n, _f(X), _f(Y), _f(M),
                                             _f(params), ctypes.byref(self.h)))

    def epoch(self, batch, lr, seed):
        loss = ctypes.c_double(0.0)
        check(self.L.ygpu_mlp_trainer_epoch(self.h, int(batch), float(lr), int(seed) & 0xFFFFFFFF, ctypes.byref(loss)))
        return loss.value

    def grad(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.zeros(self.np, dtype=np.float32)
        loss = ctypes.c_double(0.0)
        check(self.L.ygpu_mlp_trainer_grad(self.h, len(idx), _i(idx), _f(g), ctypes.byref(loss)))
        return loss.value, g

    def params(self):
        p = np.zeros(self.np, dtype=np.float32)
        check(self.L.ygpu_mlp_trainer_params(self.h, _f(p)))
        return p

    def close(self):
        if self.h:
            check(self.L.ygpu_mlp_trainer_destroy(self.h))
            self.h = vp()
