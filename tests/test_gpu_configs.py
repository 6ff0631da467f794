"""The large workloads end to end on the GPU (every stage on the device, FP64
unless stated): the C54H18 flake in FP64 against the mixed-precision mode,
the learned guess on the held-out C30H62 alkane, and the C96H24 flake with
the mixed-precision mode plus the learned guess against the plain FP64 path.

The CPU oracle is too slow for these sizes, so the FP64 GPU run is the
reference (its parity with the oracle is pinned on water and C10H22 in
test_gpu_parity.py) and convergence is also checked by the commutator
X^T (F D S - S D F) X of the converged SCF, which vanishes at the solution
whatever the arithmetic that led there."""
import os
import sys

import numpy as np
import pytest

from conftest import PKG, xyz_path

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ygpu():
    import ygpu as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


def weights(family):
    path = os.path.join(PKG, "data", "guess_mlp_%s.bin" % family)
    assert os.path.exists(path), "trained weights missing: run guess_train.py --family %s" % family
    return path


def commutator(r):
    """max |X^T (F D S - S D F) X| of a converged run (D its final density, F the FP64 Fock
    matrix built from it)."""
    F, D, S, X = r["F"], r["D"], r["S"], r["X"]
    FDS = F @ D @ S
    return float(np.abs(X.T @ (FDS - FDS.T) @ X).max())


def describe(tag, r):
    log("  %s: %.3f s (setup %.3f, DF %.3f, SCF %.3f), %d iterations%s, E = %.10f, converged %d%s"
        % (tag, r["seconds"], r["setup_s"], r["df_s"], r["scf_s"], r["iters"],
           (", FP64 from %d" % r["switch"]) if r["switch"] >= 0 else "", r["energy"], r["converged"],
           (", max|[F,D]| %.2e" % commutator(r)) if "F" in r else ""))


def test_c54h18_fp64_against_mixed(ygpu):
    """C54H18 (nbf 900, naux 3834, 171 occupied), SAD guess, |dE| < 1e-9 between FP64 energies:
    the mixed-precision run converges to the FP64 run's energy within 1e-8 Eh; one mixed Fock
    build on the converged density differs from the FP64 build by the split arithmetic's error
    only (bounds in units of the largest exchange element, K_ii <= max_i K_ii)."""
    log("case: C54H18 FP64 vs mixed")
    xyz = xyz_path("c54h18")
    fp = ygpu.solve(xyz, ene_tol=1e-9, quirk_a2=0, fock_check=True)
    describe("FP64", fp)
    mx = ygpu.solve(xyz, ene_tol=1e-9, quirk_a2=0, mixed=True, final_fock=True)
    describe("mixed", mx)
    fc = fp["fock_check"]
    log("  mixed Fock build at the converged density: max|dF| %.3e, max|d eps| %.3e (occupied %.3e)"
        % (fc["max_abs_dF"], fc["max_abs_d_orbital_energy"], fc["max_abs_d_occupied_orbital_energy"]))
    assert fp["converged"] and mx["converged"]
    assert fp["nbf"] == 900 and fp["naux"] == 3834 and fp["n_occ"] == 171
    assert mx["switch"] >= 3
    dE = mx["energy"] - fp["energy"]
    log("  dE mixed - FP64 = %.3e Eh" % dE)
    assert abs(dE) < 1e-8
    # both at the SCF solution
    assert commutator(fp) < 1e-4 and commutator(mx) < 1e-4
    kmax = 3.6    # largest K_ii of the carbon 1s shells
    assert fc["max_abs_dF"] < 2.0 ** -14 * kmax
    assert fc["max_abs_d_occupied_orbital_energy"] < 2.0 ** -14 * kmax


def test_c30h62_learned_guess_held_out(ygpu):
    """The alkane network (trained on C4-C20) on the held-out C30H62 workload (nbf 760, naux
    2910): fewer FP64 iterations than from SAD, the same converged energy."""
    log("case: C30H62 learned guess")
    xyz = xyz_path("c30h62")
    sad = ygpu.solve(xyz, ene_tol=1e-9, quirk_a2=0)
    describe("SAD", sad)
    ml = ygpu.solve(xyz, ene_tol=1e-9, quirk_a2=0, weights=weights("alkane"))
    describe("learned", ml)
    assert sad["nbf"] == 760 and sad["naux"] == 2910
    assert sad["converged"] and ml["converged"] and ml["guess_ml"] == 1
    log("  iterations saved %d; dE %.3e Eh" % (sad["iters"] - ml["iters"], ml["energy"] - sad["energy"]))
    assert ml["iters"] < sad["iters"]
    assert abs(ml["energy"] - sad["energy"]) < 1e-8


def test_c96h24_mixed_learned_against_fp64(ygpu):
    """C96H24 (nbf 1560, naux 6696, 300 occupied), |dE| < 1e-8 between FP64 energies: the
    mixed-precision run from the learned guess (flake network; C96H24 is not in its training
    set) reaches the plain FP64 run's energy within 1 uEh, in fewer iterations and less time."""
    log("case: C96H24 mixed + learned vs FP64")
    xyz = xyz_path("c96h24")
    fp = ygpu.solve(xyz, ene_tol=1e-8, quirk_a2=0, final_fock=True)
    describe("FP64 (SAD)", fp)
    mx = ygpu.solve(xyz, ene_tol=1e-8, quirk_a2=0, mixed=True, weights=weights("flake"), final_fock=True)
    describe("mixed + learned", mx)
    assert fp["nbf"] == 1560 and fp["naux"] == 6696 and fp["n_occ"] == 300
    assert fp["converged"] and mx["converged"] and mx["guess_ml"] == 1
    dE = mx["energy"] - fp["energy"]
    log("  dE %.3e Eh; speedup %.2fx" % (dE, fp["seconds"] / mx["seconds"]))
    assert abs(dE) < 1e-6
    assert commutator(fp) < 1e-4 and commutator(mx) < 1e-4
    assert mx["iters"] < fp["iters"]
    assert mx["seconds"] < fp["seconds"]
