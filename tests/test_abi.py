"""The C-ABI library loads and exports every entry point include/*.h
declares; without a device the entry points report errors instead of
crashing.  No compute call is made here."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(ygpu_\w+)\s*\(", txt))
    return names


def test_header_declares_the_api():
    d = declared()
    for n in ("ygpu_create", "ygpu_destroy", "ygpu_build_df_tensor", "ygpu_build_fock", "ygpu_build_density",
              "ygpu_set_sparsity", "ygpu_schwarz", "ygpu_ovlp_hcore", "ygpu_last_error"):
        assert n in d


def test_library_exports_every_declared_symbol():
    import ygpu
    L = ygpu.gpu_lib()
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_host_engine_exports_reference_api():
    import ygpu
    L = ygpu.host_lib()
    for n in ("init_TinySCF", "free_TinySCF", "TinySCF_init_batch_dgemm_arrays", "TinySCF_free_batch_dgemm_arrays",
              "TinySCF_compute_Hcore_Ovlp_mat", "TinySCF_compute_sq_Schwarz_scrvals", "TinySCF_prepare_sparsity",
              "TinySCF_get_initial_guess", "TinySCF_do_SCF", "TinySCF_build_DF_tensor", "TinySCF_build_FockMat",
              "TinySCF_build_DenMat", "quickSort_eigval", "TinySCF_DIIS", "TinySCF_D2Cocc"):
        assert hasattr(L, n), n


def test_errors_without_handle():
    import ygpu
    L = ygpu.gpu_lib()
    assert L.ygpu_build_fock(None, 1, None, None, None, None, None, None, None) == 1
    assert b"null context" in L.ygpu_last_error()
    assert L.ygpu_set_fock_precision(None, 1) == 1
    assert L.ygpu_build_fock_noalloc(None, 1, None, None, None, None, None, None) == 1
    assert b"null context" in L.ygpu_last_error()
    assert L.ygpu_diis_error(None, 0, None) == 1
    assert L.ygpu_destroy(None) == 0
    # page-locked host memory: a null result pointer is rejected, freeing NULL is a no-op
    assert L.ygpu_host_alloc(ctypes.c_size_t(64), None) == 1
    assert b"null result pointer" in L.ygpu_last_error()
    assert L.ygpu_host_free(None) == 0


def test_quicksort_eigval_carries_indices():
    import numpy as np
    import ygpu
    L = ygpu.host_lib()
    rng = np.random.RandomState(1)
    ev = rng.randn(50)
    ev[7] = ev[8] = ev[30]           # exact ties
    idx = np.arange(50, dtype=np.int32)
    orig = ev.copy()
    L.quickSort_eigval(ev.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 0, 49)
    assert np.all(np.diff(ev) >= 0)
    assert sorted(idx.tolist()) == list(range(50))        # a permutation (no duplicated index)
    assert np.array_equal(orig[idx], ev)


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") is None and
                    not os.path.exists("/dev/kfd"), reason="runs only where no device node exists")
def test_create_without_device_reports_nodevice():
    import ygpu
    if ygpu.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(ygpu.GpuError) as e:
        ygpu.Engine(0)
    assert e.value.status == 6
