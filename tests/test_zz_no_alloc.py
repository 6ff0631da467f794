"""ygpu_build_fock_noalloc (the FP64 Fock build with its device buffers in the
library's static scratch array) against ygpu_build_fock on the same inputs,
through the C-ABI: C10H22 with its 41 occupied orbitals, then the same engine
with a smaller occupied block (17 columns), then a second engine on water
(nbf 25), so each later call runs on the scratch contents the earlier one
left.  Bar: max |dF|, |dJ|, |dK| < 1e-12 (the same kernels in the same launch
shapes; the results are expected to be bit-identical)."""
import sys

import numpy as np
import pytest

from conftest import system

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ygpu():
    import ygpu as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


def oracle_inputs(name):
    import oracle
    s = system(name)
    o = oracle.OracleSCF(s, niters=60)
    o.setup()
    o.build_df_tensor()
    n = s.nbf
    return s, o, o.mat("D_mat").copy(), o.d("Cocc_mat", (n, s.n_occ)).copy(), o.mat("Hcore_mat").copy()


def compare(tag, a, b):
    err = {k: float(np.max(np.abs(a[k] - b[k]))) for k in ("F", "J", "K")}
    same = all(np.array_equal(a[k], b[k]) for k in ("F", "J", "K"))
    log("  %s: max|dF| %.2e max|dJ| %.2e max|dK| %.2e bit-identical %s" % (tag, err["F"], err["J"], err["K"], same))
    assert max(err.values()) < 1e-12, (tag, err)


def test_noalloc_fock_matches_the_allocating_entry_point(ygpu):
    log("case: no-allocation Fock build, c10h22 then a smaller occupied block, then water")
    s, o, D, C, H = oracle_inputs("c10h22")
    e = ygpu.Engine(0)
    try:
        e.set_system(s)
        e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                       o.d("df_sp_scrval"))
        e.build_df_tensor()
        compare("c10h22 n_occ 41", e.build_fock_noalloc(D, C, H), e.build_fock(D, C, H))
        # a smaller input on the same scratch: 17 occupied columns (D from the same columns)
        C17 = np.ascontiguousarray(C[:, :17])
        D17 = C17 @ C17.T
        ref = e.build_fock(D17, C17, H)
        got = e.build_fock_noalloc(D17, C17, H)
        compare("c10h22 n_occ 17", got, ref)
        # the mixed mode is refused with a clear error, and the engine stays usable
        e.set_fock_precision(ygpu.FOCK_MIXED)
        with pytest.raises(ygpu.GpuError) as err:
            e.build_fock_noalloc(D, C, H)
        assert "FP64 arithmetic only" in str(err.value)
        e.set_fock_precision(ygpu.FOCK_FP64)
    finally:
        e.close()
    s, o, D, C, H = oracle_inputs("water")
    e = ygpu.Engine(0)
    try:
        e.set_system(s)
        e.set_sparsity(o.i("bf_pair_mask"), o.i("uniq_sp_lid"), o.i("uniq_sp_rid"), o.d("sp_scrval"),
                       o.d("df_sp_scrval"))
        e.build_df_tensor()
        got = e.build_fock_noalloc(D, C, H)
        compare("water", got, e.build_fock(D, C, H))
        # and against the oracle's own Fock matrix (the reference algorithm on the CPU)
        o.build_fock()
        errF = float(np.max(np.abs(got["F"] - o.mat("F_mat"))))
        log("  water vs the oracle: max|dF| %.2e" % errF)
        assert errF < 1e-10
    finally:
        e.close()


def test_noalloc_rejects_a_workload_beyond_the_scratch(ygpu):
    """C54H18's X needs gigabytes: a clear error, nothing launched."""
    log("case: no-allocation Fock build beyond the scratch size")
    import basis  # noqa: F401
    s = system("c54h18")
    e = ygpu.Engine(0)
    try:
        e.set_system(s)
        n = s.nbf
        # only the diagonal pairs kept (a small tensor to upload): X alone is still
        # nbf * n_occ * naux_pad * 8 = 4.7 GB, and the size check precedes any launch
        mask = np.full((n, n), -1, dtype=np.int32)
        mask[np.arange(n), np.arange(n)] = np.arange(n)
        e.set_sparsity(mask, np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32),
                       np.ones(s.basis["nshells"] ** 2), np.ones(s.auxbasis["nshells"]))
        e.set_df_tensor(np.zeros((n, s.naux)))
        D = np.zeros((n, n))
        C = np.zeros((n, s.n_occ))
        with pytest.raises(ygpu.GpuError) as err:
            e.build_fock_noalloc(D, C, D)
        log("  %s" % err.value)
        assert "scratch" in str(err.value)
    finally:
        e.close()
