"""Spherical-harmonic basis option on the GPU (YATSCF_BASIS_TYPE=spherical, or a .gbs file
whose first token is "spherical").  No reference counterpart exists (its Simint path is
Cartesian), so the results are pinned by invariants:
  - function counts: water 24 / 84, C10H22 250 / 868 (orbital / auxiliary), the sizes quoted
    for cc-pVDZ / cc-pVDZ-RI;
  - rotational invariance: the span of the solid harmonics of each shell is invariant under
    rotations of the molecule only if every transform column is right, so the converged
    energy of a rotated molecule must equal the unrotated one (a wrong coefficient leaves a
    rotation-dependent subspace and energy);
  - variational bound: with the auxiliary basis Cartesian in both runs, the spherical orbital
    basis is a subspace of the Cartesian one, so its converged energy is not lower;
  - s/p shells only (d and f shells removed from both basis files): T is the identity and the
    spherical run is the Cartesian run;
  - the mixed-precision SCF in the spherical basis reaches the FP64 spherical energy."""
import os
import shutil
import sys
import tempfile

import numpy as np
import pytest

from conftest import PKG, xyz_path

pytestmark = pytest.mark.gpu

CCPVDZ = os.path.join(PKG, "data", "cc-pvdz", "cc-pvdz.gbs")
RI = os.path.join(PKG, "data", "cc-pVDZ-RI", "cc-pVDZ-RI.gbs")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ygpu():
    import ygpu as m
    assert m.device_count() >= 1, "no GPU visible"
    return m


def read_xyz(path):
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    el, xyz = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        el.append(t[0])
        xyz.append([float(v) for v in t[1:4]])
    return lines[1], el, np.array(xyz)


def write_xyz(path, charge, el, xyz):
    with open(path, "w") as f:
        f.write("%d\n%s\n" % (len(el), charge))
        for e, (x, y, z) in zip(el, xyz):
            f.write("%s %.12f %.12f %.12f\n" % (e, x, y, z))


def rotated(src, dst, seed):
    q, _ = np.linalg.qr(np.random.RandomState(seed).randn(3, 3))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    charge, el, xyz = read_xyz(src)
    write_xyz(dst, charge, el, (xyz - xyz.mean(0)) @ q.T + np.array([0.3, -0.2, 0.1]))


def gbs_variant(src, dst_dir, first_line=None, max_am=None):
    """copy of a .gbs (and its SAD files) with an optional first line and without shells above
    max_am"""
    d = os.path.dirname(src)
    os.makedirs(dst_dir, exist_ok=True)
    for f in os.listdir(d):
        if f.endswith(".dat"):
            shutil.copy(os.path.join(d, f), dst_dir)
    out, skip = [], False
    for ln in open(src).read().splitlines():
        t = ln.split()
        if len(t) == 3 and t[0] in ("S", "P", "D", "F", "G", "SP"):        # "<shell> <nexp> <scale>"
            am = "SPDFG".index(t[0][-1])
            skip = max_am is not None and am > max_am
            if skip:
                continue
        elif t and t[0][0].isalpha() or (t and t[0].startswith("*")):
            skip = False
        elif skip:
            continue
        out.append(ln)
    if first_line is not None:
        out.insert(0, first_line)
    dst = os.path.join(dst_dir, os.path.basename(src))
    open(dst, "w").write("\n".join(out) + "\n")
    return dst


def run(ygpu, xyz, **kw):
    kw.setdefault("ene_tol", 1e-11)
    kw.setdefault("quirk_a2", 0)
    return ygpu.solve(xyz, niters=80, **kw)


@pytest.mark.parametrize("name,nbf,naux", [("water", 24, 84), ("c10h22", 250, 868)])
def test_function_counts_and_rotation_invariance(ygpu, name, nbf, naux):
    log("case: spherical counts and rotation invariance", name)
    src = xyz_path(name) if name != "water" else os.path.join(PKG, "data", "water.xyz")
    with tempfile.TemporaryDirectory() as tmp:
        rot = os.path.join(tmp, "rot.xyz")
        rotated(src, rot, 11)
        a = run(ygpu, src, basis_type="spherical")
        b = run(ygpu, rot, basis_type="spherical")
        c = run(ygpu, src, basis_type="cartesian")
        d = run(ygpu, rot, basis_type="cartesian")
    log("  spherical: nbf %d naux %d E %.12f rotated %.12f (dE %.2e); Cartesian E %.12f rotated %.12f "
        "(dE %.2e); spherical - Cartesian %.3e"
        % (a["nbf_sph"], a["naux"], a["energy"], b["energy"], b["energy"] - a["energy"], c["energy"], d["energy"],
           d["energy"] - c["energy"], a["energy"] - c["energy"]))
    assert a["converged"] and b["converged"] and c["converged"] and d["converged"]
    assert (a["nbf_sph"], a["naux"]) == (nbf, naux)
    assert a["sph_orb"] == 1 and a["sph_aux"] == 1 and c["sph_orb"] == 0
    assert abs(d["energy"] - c["energy"]) < 1e-9       # the Cartesian control
    assert abs(b["energy"] - a["energy"]) < 1e-9
    # the spherical basis drops the x^2+y^2+z^2 components: a different, higher energy
    assert a["energy"] - c["energy"] > 1e-4


def test_spherical_orbital_basis_is_variational(ygpu):
    log("case: spherical orbital basis with a Cartesian auxiliary basis, water")
    xyz = os.path.join(PKG, "data", "water.xyz")
    with tempfile.TemporaryDirectory() as tmp:
        sph = gbs_variant(CCPVDZ, os.path.join(tmp, "sph"), first_line="spherical")
        a = run(ygpu, xyz, basis=sph, auxbasis=RI)                  # file flags: spherical / Cartesian
        c = run(ygpu, xyz, basis_type="cartesian")
    log("  E spherical orbitals %.12f (nbf %d, naux %d), Cartesian %.12f (naux %d), difference %.3e"
        % (a["energy"], a["nbf_sph"], a["naux"], c["energy"], c["naux"], a["energy"] - c["energy"]))
    assert a["sph_orb"] == 1 and a["sph_aux"] == 0 and a["naux"] == c["naux"]
    assert a["energy"] >= c["energy"] - 1e-10
    assert a["energy"] - c["energy"] > 1e-4


def test_sp_shells_only_spherical_is_cartesian(ygpu):
    log("case: s/p shells only, spherical against Cartesian, water")
    xyz = os.path.join(PKG, "data", "water.xyz")
    with tempfile.TemporaryDirectory() as tmp:
        bs = gbs_variant(CCPVDZ, os.path.join(tmp, "o"), max_am=1)
        ab = gbs_variant(RI, os.path.join(tmp, "a"), max_am=1)
        a = run(ygpu, xyz, basis=bs, auxbasis=ab, basis_type="spherical")
        c = run(ygpu, xyz, basis=bs, auxbasis=ab, basis_type="cartesian")
    log("  E spherical %.14f Cartesian %.14f (nbf %d / %d, naux %d / %d)"
        % (a["energy"], c["energy"], a["nbf_sph"], c["nbf"], a["naux"], c["naux"]))
    assert a["nbf_sph"] == c["nbf"] and a["naux"] == c["naux"]
    assert abs(a["energy"] - c["energy"]) < 1e-11


def test_mixed_precision_in_the_spherical_basis(ygpu):
    log("case: mixed-precision SCF, spherical basis, c10h22")
    xyz = xyz_path("c10h22")
    f = run(ygpu, xyz, basis_type="spherical")
    m = run(ygpu, xyz, basis_type="spherical", mixed=True)
    log("  E FP64 %.12f mixed %.12f dE %.2e iterations %d / %d" % (f["energy"], m["energy"], m["energy"] - f["energy"],
                                                                    f["iters"], m["iters"]))
    assert f["converged"] and m["converged"]
    assert abs(m["energy"] - f["energy"]) < 1e-8
