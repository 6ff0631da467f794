"""The drop-in boundary: the three hot-path files that replace the reference's
build_DF_tensor.c, build_Fock.c and build_density.c read only the fields of
the reference's TinySCF struct (TinySCF.h:12-100) plus the device handle
`gpu`.  The struct declared in host/TinySCF.h is the reference's field list,
then `gpu` in place of the Simint handle, then this engine's additions; the
tests cut the additions away, compile the three files against what is left
(CPU), and on the GPU drive them through a small C shim on that reduced
struct -- the situation of a maintainer who drops the files into the
reference tree -- and compare with the device engine's API."""
import ctypes
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from conftest import PKG, ROOT, system

HOST = os.path.join(PKG, "host")
DROPIN = ("build_DF_tensor.c", "build_Fock.c", "build_density.c")


def reduced_tree(tmp):
    """host headers and the three drop-in files in tmp, with TinySCF.h cut to the reference's
    fields plus `gpu` (everything from the additions marker to the end of the struct removed)."""
    src = open(os.path.join(HOST, "TinySCF.h")).read()
    a = src.index("    /* additions */")
    b = src.index("};", a)
    red = src[:a] + src[b:]
    assert "void   *gpu;" in red and "int    quirk_a5;" not in red and "int    mixed;" not in red
    with open(os.path.join(tmp, "TinySCF.h"), "w") as f:
        f.write(red)
    for name in DROPIN + ("build_DF_tensor.h", "build_Fock.h", "build_density.h", "utils.h", "CMS.h"):
        shutil.copy(os.path.join(HOST, name), tmp)
    return tmp


def compile_dropin(tmp, out=None, extra=()):
    cmd = ["gcc", "-std=gnu11", "-Wall", "-Werror", "-fPIC", "-I", tmp, "-I", os.path.join(ROOT, "include")]
    srcs = [os.path.join(tmp, n) for n in DROPIN] + list(extra)
    if out is None:
        cmd += ["-fsyntax-only"] + srcs
    else:
        cmd += ["-shared", "-o", out] + srcs + [os.path.join(HOST, "utils.c"), "-L", PKG, "-lyatscf_gpu",
                                                 "-Wl,-rpath," + PKG, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_files_compile_against_the_reference_fields():
    with tempfile.TemporaryDirectory() as tmp:
        compile_dropin(reduced_tree(tmp))


def test_dropin_files_read_no_engine_addition():
    """Belt and braces: none of the three files names a field the reference's struct lacks."""
    src = open(os.path.join(HOST, "TinySCF.h")).read()
    a = src.index("    /* additions */")
    b = src.index("};", a)
    import re
    extra = set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*(?:\[[^\]]*\])?\s*;", src[a:b]))
    for name in DROPIN:
        body = open(os.path.join(HOST, name)).read()
        used = set(re.findall(r"T->([A-Za-z_][A-Za-z0-9_]*)", body))
        assert not (used & extra), (name, used & extra)


SHIM = r"""
#include <string.h>
#include "TinySCF.h"
#include "build_DF_tensor.h"
#include "build_Fock.h"
#include "build_density.h"

/* the reference's struct, filled as its init and setup would leave it, then the three
 * hot-path calls in the reference's order (main.c:40, TinySCF.c:395, :415) */
int shim_run(void *gpu, int nbf, int n_occ, int naux, double *Hcore, double *X, double *D, double *Cocc,
             double *J, double *K, double *F, double *Jpq, double *Fo, double *D1, double *C1, double *eig,
             int *ev_idx)
{
    struct TinySCF_struct T;
    memset(&T, 0, sizeof(T));
    T.gpu = gpu;
    T.nbasfuncs = nbf;
    T.n_occ = n_occ;
    T.df_nbf = naux;
    T.mat_size = nbf * nbf;
    T.Hcore_mat = Hcore;
    T.X_mat = X;
    T.D_mat = D;
    T.Cocc_mat = Cocc;
    T.J_mat = J;
    T.K_mat = K;
    T.F_mat = F;
    T.Jpq = Jpq;
    T.eigval = eig;
    T.ev_idx = ev_idx;
    TinySCF_build_DF_tensor(&T);
    TinySCF_build_FockMat(&T);
    /* the DIIS step leaves X^T F X in F_mat (DIIS.c:30-40); here the caller's, in its own buffer
     * so that the Fock matrix above stays readable */
    T.F_mat = Fo;
    T.D_mat = D1;
    T.Cocc_mat = C1;
    TinySCF_build_DenMat(&T);
    return 0;
}
"""


@pytest.mark.gpu
def test_dropin_files_run_on_the_reference_fields():
    """Water: the three files on the reduced struct (host Hcore and X passed in, no device
    caches set up beforehand) give the Fock matrix and density the engine API gives."""
    import oracle
    import ygpu
    print("case: drop-in files on the reference's struct fields, water", flush=True)
    st = system("water")
    o = oracle.OracleSCF(st, niters=10)
    o.setup()
    o.build_df_tensor()
    n, no, na = st.nbf, st.n_occ, st.naux
    D, C = o.mat("D_mat").copy(), o.d("Cocc_mat", (n, no)).copy()
    H, X = o.mat("Hcore_mat").copy(), o.mat("X_mat").copy()
    o.build_fock()
    F_ref = o.mat("F_mat").copy()
    with tempfile.TemporaryDirectory() as tmp:
        reduced_tree(tmp)
        with open(os.path.join(tmp, "shim.c"), "w") as f:
            f.write(SHIM)
        lib = os.path.join(tmp, "libdropin.so")
        compile_dropin(tmp, lib, extra=[os.path.join(tmp, "shim.c")])
        L = ctypes.CDLL(lib)
        e = ygpu.Engine(0)
        try:
            e.set_system(st)
            e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                           o.d("df_sp_scrval"))
            arrs = dict(J=np.zeros((n, n)), K=np.zeros((n, n)), F=np.zeros((n, n)), Jpq=np.zeros((na, na)),
                        D1=np.zeros((n, n)), C1=np.zeros((n, no)), eig=np.zeros(n))
            Fo = X.T @ F_ref @ X
            idx = np.zeros(n, dtype=np.int32)
            dp = ctypes.POINTER(ctypes.c_double)
            c = lambda a: a.ctypes.data_as(dp)   # noqa: E731
            L.shim_run.restype = ctypes.c_int
            assert L.shim_run(e.h, n, no, na, c(H), c(X), c(D), c(C), c(arrs["J"]), c(arrs["K"]), c(arrs["F"]),
                              c(arrs["Jpq"]), c(Fo), c(arrs["D1"]), c(arrs["C1"]), c(arrs["eig"]),
                              idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
            D1, C1, ev, _ = e.build_density(Fo, X, no)
        finally:
            e.close()
    errF = float(np.abs(arrs["F"] - F_ref).max())
    errD = float(np.abs(arrs["D1"] - D1).max())
    print("  max|F - F_oracle| %.2e, max|D - D_api| %.2e" % (errF, errD), flush=True)
    sys.stdout.flush()
    assert errF < 1e-10
    assert np.array_equal(arrs["D1"], D1) and np.array_equal(arrs["eig"], ev)
    assert np.array_equal(idx, np.arange(n))


@pytest.mark.gpu
def test_one_sided_pair_mask_is_taken_as_the_union():
    """A mask that keeps (i, j) but not (j, i) -- what the reference's one-orientation Schwarz
    index can produce (TinySCF.c:153, quirk A5) -- is accepted and read as the union of both
    orientations: the same Fock matrix as the symmetric mask, bit for bit."""
    import oracle
    import ygpu
    print("case: one-sided pair mask, c10h22", flush=True)
    st = system("c10h22")
    o = oracle.OracleSCF(st, niters=10)
    o.setup()
    n, no = st.nbf, st.n_occ
    mask = o.i("bf_pair_mask").reshape(n, n).copy()
    D, C, H = o.mat("D_mat").copy(), o.d("Cocc_mat", (n, no)).copy(), o.mat("Hcore_mat").copy()
    iu = np.argwhere(np.triu(mask >= 0, 1))          # kept pairs with i < j
    one = mask.copy()
    for i, j in iu[::7]:                      # drop the lower orientation of every 7th kept pair
        one[j, i] = -1
    assert np.any((one >= 0) != (one.T >= 0))
    out = []
    for m in (mask, one):
        e = ygpu.Engine(0)
        try:
            e.set_system(st)
            e.set_sparsity(m, o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"), o.d("df_sp_scrval"))
            assert e.layout_info()[1] == int(np.sum(mask >= 0))
            e.build_df_tensor()
            out.append(e.build_fock(D, C, H)["F"])
        finally:
            e.close()
    assert np.array_equal(out[0], out[1])
