"""Writes tests/golden/integrals_mpmath.json: integral values of random small
contracted shells computed with the arbitrary-precision Obara-Saika code of
tests/os_mpmath.py (an independent formulation from the oracle's
McMurchie-Davidson code).  Inputs and outputs only; rerun to regenerate:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import os_mpmath as OS  # noqa: E402

rng = np.random.RandomState(20261203)


def shell(am, nprim):
    return dict(am=am, xyz=rng.uniform(-1.5, 1.5, 3).tolist(), exps=rng.uniform(0.15, 4.0, nprim).tolist(),
                coefs=rng.uniform(0.2, 1.0, nprim).tolist())


def t(s):
    return (s["am"], s["xyz"], s["exps"], s["coefs"])


out = dict(eri3=[], eri2=[], one=[], eri4=[], boys=[])
for la in range(3):
    for lb in range(la + 1):
        for lc in range(4):
            a, b, c = shell(la, 2 if la < 2 else 1), shell(lb, 2), shell(lc, 1)
            out["eri3"].append(dict(a=a, b=b, c=c, val=[float(v) for v in OS.eri3(t(a), t(b), t(c))]))
for lc in range(4):
    for ld in range(4):
        c, d = shell(lc, 1), shell(ld, 2)
        out["eri2"].append(dict(c=c, d=d, val=[float(v) for v in OS.eri2(t(c), t(d))]))
Zs = [6.0, 1.0, 8.0]
cen = rng.uniform(-2, 2, (3, 3)).tolist()
for la in range(3):
    for lb in range(3):
        a, b = shell(la, 2), shell(lb, 3)
        out["one"].append(dict(a=a, b=b, Z=Zs, centres=cen,
                               S=[[float(v) for v in r] for r in OS.overlap(t(a), t(b))],
                               T=[[float(v) for v in r] for r in OS.kinetic(t(a), t(b))],
                               V=[[float(v) for v in r] for r in OS.nuclear(t(a), t(b), Zs, cen)]))
for la in range(3):
    for lb in range(3):
        a, b = shell(la, 2), shell(lb, 1)
        na, nb = (la + 1) * (la + 2) // 2, (lb + 1) * (lb + 2) // 2
        for _ in range(3):
            ia, ib, ic, id_ = [int(rng.randint(0, n)) for n in (na, nb, na, nb)]
            out["eri4"].append(dict(a=a, b=b, idx=[ia, ib, ic, id_],
                                    val=float(OS.eri4(t(a), t(b), t(a), t(b), ia, ib, ic, id_))))
for T in [0.0, 1e-9, 1e-3, 0.05, 0.3, 1.0, 3.3, 7.77, 12.5, 25.0, 33.3, 39.99, 40.0, 40.02, 55.5, 120.0, 900.0]:
    out["boys"].append(dict(T=T, F=[float(OS.boys(m, T)) for m in range(17)]))
with open(os.path.join(HERE, "integrals_mpmath.json"), "w") as f:
    json.dump(out, f, indent=0)
print("wrote", sum(len(v) for v in out.values()), "records")
