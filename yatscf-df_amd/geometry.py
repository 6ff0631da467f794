"""Seeded molecular geometries for the DF-RHF workloads.

All coordinates are in Angstrom and are written as ``.xyz`` files in the
format the reference reads (atom count, charge line, ``El x y z``;
``CMS_BasisSet.c:310-426``).

* ``water()``        O at the origin, O-H 0.9572 A, H-O-H 104.52 deg.
* ``alkane(n)``      all-trans zig-zag C_nH_(2n+2), C-C 1.54 A, C-H 1.09 A,
                     tetrahedral angles, optional torsions and uniform jitter.
* ``graphene(r)``    hexagonal-lattice flake of all rings within hexagonal
                     distance r of a central ring (r=2: circumcoronene C54H18,
                     r=3: C96H24), C-C 1.42 A, C-H 1.09 A, planar.
"""
import math

import numpy as np

TETRA = math.acos(-1.0 / 3.0)          # 109.4712 deg


def _unit(v):
    return v / np.linalg.norm(v)


def water():
    half = math.radians(104.52) / 2.0
    r = 0.9572
    return ["O", "H", "H"], np.array([[0.0, 0.0, 0.0],
                                      [r * math.sin(half), 0.0, r * math.cos(half)],
                                      [-r * math.sin(half), 0.0, r * math.cos(half)]])


def _rotate_about(v, axis, ang):
    axis = _unit(axis)
    return (v * math.cos(ang) + np.cross(axis, v) * math.sin(ang)
            + axis * np.dot(axis, v) * (1.0 - math.cos(ang)))


def alkane(n, jitter=0.0, seed=None, torsion=0.0):
    """All-trans C_nH_(2n+2).  ``torsion`` (deg) is the half-width of uniform
    random rotations applied about each C-C bond to everything beyond it."""
    cc, ch = 1.54, 1.09
    half = TETRA / 2.0
    C = []
    for k in range(n):
        C.append([k * cc * math.sin(half), (0.5 if k % 2 == 0 else -0.5) * cc * math.cos(half), 0.0])
    C = np.array(C)
    syms = ["C"] * n
    H = []
    zax = np.array([0.0, 0.0, 1.0])
    for k in range(n):
        nbr = [C[j] - C[k] for j in (k - 1, k + 1) if 0 <= j < n]
        if len(nbr) == 2:
            u = _unit(-(_unit(nbr[0]) + _unit(nbr[1])))
            for sgn in (1.0, -1.0):
                H.append(C[k] + ch * (u * math.cos(half) + sgn * zax * math.sin(half)))
        else:
            if n == 1:
                dirs = [np.array([1, 1, 1]), np.array([1, -1, -1]), np.array([-1, 1, -1]), np.array([-1, -1, 1])]
                for d in dirs:
                    H.append(C[k] + ch * _unit(d.astype(float)))
                continue
            b = _unit(nbr[0])
            e1 = _unit(np.cross(b, zax))
            e2 = np.cross(b, e1)
            # staggered methyl: one H in the chain plane, anti to the C-C bond's neighbour
            base = C[k] - C[k + (1 if k == 0 else -1)]
            for i in range(3):
                phi = math.radians(90.0 + 120.0 * i)
                d = math.cos(TETRA) * b + math.sin(TETRA) * (math.cos(phi) * e1 + math.sin(phi) * e2)
                H.append(C[k] + ch * d)
            del base
    H = np.array(H)
    xyz = np.vstack([C, H])
    syms = syms + ["H"] * len(H)
    # attach each H to its carbon (nearest C) for torsion moves
    owner = np.argmin(np.linalg.norm(H[:, None, :] - C[None, :, :], axis=2), axis=1)
    rng = np.random.RandomState(seed) if seed is not None else None
    if torsion and rng is not None and n >= 3:
        for k in range(1, n - 2):
            ang = math.radians(rng.uniform(-torsion, torsion))
            axis = C[k + 1] - C[k]
            origin = C[k + 1]
            mask_c = np.arange(n) > k + 1
            mask_h = owner > k + 1
            C[mask_c] = np.array([origin + _rotate_about(p - origin, axis, ang) for p in C[mask_c]]).reshape(-1, 3)
            H[mask_h] = np.array([origin + _rotate_about(p - origin, axis, ang) for p in H[mask_h]]).reshape(-1, 3)
        xyz = np.vstack([C, H])
    if jitter and rng is None:
        rng = np.random.RandomState(0)
    if jitter:
        xyz = xyz + rng.uniform(-jitter, jitter, size=xyz.shape)
    return syms, xyz


def graphene(rings, jitter=0.0, seed=None):
    a, ch = 1.42, 1.09
    s3 = math.sqrt(3.0)
    a1 = np.array([s3 * a, 0.0])
    a2 = np.array([s3 * a / 2.0, 1.5 * a])
    pts = {}
    for i in range(-rings, rings + 1):
        for j in range(-rings, rings + 1):
            if max(abs(i), abs(j), abs(i + j)) > rings:
                continue
            c = i * a1 + j * a2
            for k in range(6):
                ang = math.radians(30.0 + 60.0 * k)
                p = c + a * np.array([math.cos(ang), math.sin(ang)])
                pts[(round(p[0], 4), round(p[1], 4))] = p
    C = np.array(sorted(pts.values(), key=lambda p: (round(p[1], 3), p[0])))
    H = []
    for k in range(len(C)):
        d = np.linalg.norm(C - C[k], axis=1)
        nb = np.nonzero((d > 0.1) & (d < a * 1.1))[0]
        if len(nb) == 2:
            u = _unit(2 * C[k] - C[nb[0]] - C[nb[1]])
            H.append(C[k] + ch * u)
    H = np.array(H)
    xyz = np.zeros((len(C) + len(H), 3))
    xyz[:len(C), :2] = C
    xyz[len(C):, :2] = H
    syms = ["C"] * len(C) + ["H"] * len(H)
    if jitter:
        rng = np.random.RandomState(seed)
        xyz = xyz + rng.uniform(-jitter, jitter, size=xyz.shape)
    return syms, xyz


def xyz_text(syms, xyz, charge=0):
    lines = ["%d" % len(syms), "%d" % charge]
    for s, p in zip(syms, xyz):
        lines.append("%-2s %20.12f %20.12f %20.12f" % (s, p[0], p[1], p[2]))
    return "\n".join(lines) + "\n"


def write_xyz(path, syms, xyz, charge=0):
    with open(path, "w") as f:
        f.write(xyz_text(syms, xyz, charge))
    return path


# The named workloads (seeded; the seeds are part of the workload definition).
WORKLOADS = {
    "water": lambda: water(),
    "c10h22": lambda: alkane(10, jitter=0.02, seed=1234),
    "c54h18": lambda: graphene(2, jitter=0.01, seed=1234),
    "c30h62": lambda: alkane(30, jitter=0.05, seed=4321, torsion=30.0),
    "c96h24": lambda: graphene(3, jitter=0.01, seed=1234),
}


def workload(name):
    return WORKLOADS[name]()
