"""Pins the oracle's integrals (McMurchie-Davidson, oracle/ora_ints.c)
against values from an independent arbitrary-precision Obara-Saika code
(tests/os_mpmath.py, fixtures in tests/golden/integrals_mpmath.json) and
against closed forms and invariances."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN

with open(os.path.join(GOLDEN, "integrals_mpmath.json")) as f:
    GOLD = json.load(f)


def sh(d):
    return oracle.Shell(d["am"], d["xyz"], d["exps"], d["coefs"])


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.abs(b).max(), 1e-300)


def test_boys_function():
    for rec in GOLD["boys"]:
        F = oracle.boys(16, rec["T"])
        ref = np.array(rec["F"])
        assert np.max(np.abs(F - ref) / ref) < 2e-15, rec["T"]


def test_boys_special_values():
    # F_m(0) = 1/(2m+1); F_0(T) = sqrt(pi/T)/2 erf(sqrt T)
    F = oracle.boys(10, 0.0)
    assert np.allclose(F, [1.0 / (2 * m + 1) for m in range(11)], rtol=1e-15)
    for T in (0.01, 0.7, 5.0, 22.0, 41.0, 80.0):
        assert abs(oracle.boys(0, T)[0] - 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))) < 2e-16


@pytest.mark.parametrize("k", range(24))
def test_eri3_vs_mpmath(k):
    rec = GOLD["eri3"][k]
    got = oracle.eri3(sh(rec["a"]), sh(rec["b"]), sh(rec["c"]))
    assert rel(got, rec["val"]) < 2e-14, (rec["a"]["am"], rec["b"]["am"], rec["c"]["am"])


@pytest.mark.parametrize("k", range(16))
def test_eri2_vs_mpmath(k):
    rec = GOLD["eri2"][k]
    assert rel(oracle.eri2(sh(rec["c"]), sh(rec["d"])), rec["val"]) < 2e-14


@pytest.mark.parametrize("k", range(9))
def test_one_electron_vs_mpmath(k):
    rec = GOLD["one"][k]
    A, B = sh(rec["a"]), sh(rec["b"])
    assert rel(oracle.overlap(A, B), rec["S"]) < 2e-14
    assert rel(oracle.kinetic(A, B), rec["T"]) < 2e-14
    assert rel(oracle.nuclear(A, B, rec["Z"], np.array(rec["centres"])), rec["V"]) < 5e-14


def test_eri4_elements_vs_mpmath():
    for rec in GOLD["eri4"]:
        A, B = sh(rec["a"]), sh(rec["b"])
        got = oracle.eri4(A, B, A, B, *rec["idx"])
        assert abs(got - rec["val"]) <= 2e-14 * max(abs(rec["val"]), 1e-300) + 1e-18


def test_ssss_closed_form():
    # (ss|ss) of unit-coefficient primitives: 2 pi^2.5/(pq sqrt(p+q)) K_AB K_CD F0(rho |PQ|^2)
    a, b, c = 0.9, 1.7, 0.4
    A, B, C = np.array([0.1, 0.2, -0.3]), np.array([1.0, -0.4, 0.5]), np.array([-0.7, 0.3, 1.1])
    p = a + b
    P = (a * A + b * B) / p
    K = math.exp(-a * b / p * np.sum((A - B) ** 2))
    rho = p * c / (p + c)
    T = rho * np.sum((P - C) ** 2)
    F0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    ref = 2 * math.pi ** 2.5 / (p * c * math.sqrt(p + c)) * K * F0
    got = oracle.eri3(oracle.Shell(0, A, [a], [1.0]), oracle.Shell(0, B, [b], [1.0]), oracle.Shell(0, C, [c], [1.0]))
    assert abs(got[0, 0, 0] - ref) < 1e-15 * ref


def test_eri3_symmetry_and_translation():
    rng = np.random.RandomState(5)
    A = oracle.Shell(2, rng.randn(3), [0.8], [1.0])
    B = oracle.Shell(1, rng.randn(3), [0.5, 1.9], [0.4, 0.6])
    C = oracle.Shell(3, rng.randn(3), [1.1], [1.0])
    ab = oracle.eri3(A, B, C)
    ba = oracle.eri3(B, A, C)
    assert np.max(np.abs(ab - ba.transpose(1, 0, 2))) < 1e-15 * np.abs(ab).max()
    s = np.array([0.3, -1.2, 2.0])
    A2 = oracle.Shell(2, A.s.xyz[:] + s, [0.8], [1.0])
    B2 = oracle.Shell(1, np.array(B.s.xyz[:]) + s, [0.5, 1.9], [0.4, 0.6])
    C2 = oracle.Shell(3, np.array(C.s.xyz[:]) + s, [1.1], [1.0])
    assert np.max(np.abs(oracle.eri3(A2, B2, C2) - ab)) < 1e-13 * np.abs(ab).max()


def test_normalisation_convention():
    """x^l-normalised shells: axis components have unit norm, xy of a d shell 1/3."""
    import basis
    for l in range(4):
        exps = [3.1, 0.7, 0.15]
        c = basis.normalise_shell(l, exps, [0.3, 0.5, 0.4])
        S = oracle.overlap(oracle.Shell(l, [0, 0, 0], exps, c), oracle.Shell(l, [0, 0, 0], exps, c))
        assert abs(S[0, 0] - 1.0) < 1e-14
        if l == 2:
            assert abs(S[1, 1] - 1.0 / 3.0) < 1e-14
