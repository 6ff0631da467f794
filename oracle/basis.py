"""Input parsing for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Restates libCMS's conventions so that the oracle sees the same chemical system
the reference would:

* ``.xyz``: atom count on line 1 (``CMS_BasisSet.c:327-337``), net charge =
  ``atof(line 2)`` truncated to int (``:345``), coordinates in Angstrom scaled
  by ``A2BOHR = 1/0.52917720859`` (``:39``, ``:371-373``), element symbols
  from the H..Sn table (``:44-56``), atom-count mismatch is an error
  (``:401-406``), nuclear repulsion by the closed form (``:409-421``).
* ``.gbs``: first token ``cartesian``/``spherical`` else Cartesian
  (``:468-483``), element blocks ended by ``****``, shell lines
  ``<letters> <nexp> <scale>`` with ``SP`` split into an s and a p shell that
  share exponents (``:556-616``), scale factor ignored.
* SAD guess ``<El>.dat`` next to the ``.gbs``: nfunc^2 lines, a one-sided
  symmetry test ``g[j,k]-g[k,j] > 1e-12``; a missing or bad file gives a zero
  block (``:657-746``).
* shell order: atoms in xyz order, each element's shells in file order
  (``CMS_parse_molecule``, ``:192-299``); n_occ = (Z_tot - Q) // 2
  (``:788-791``).

Normalisation: contracted Cartesian shells are normalised so the axis
component x^l has unit norm (the convention Simint applies in
``simint_normalize_shells``, ``CMS_Simint.c:144-145``; the d/f convention is
not pinned by any reference fixture, see tests/test_oracle_basis.py).
"""
import math
import os
import re

import numpy as np

A2BOHR = 1.0 / 0.52917720859
ETABLE = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P",
          "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
          "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh",
          "Pd", "Ag", "Cd", "In", "Sn"]
MTABLE = "SPDFG"


class InputError(ValueError):
    pass


def c_atof(s):
    """strtod-style prefix parse (atof returns 0 when nothing parses)."""
    s = s.lstrip()
    best = 0.0
    for n in range(len(s), 0, -1):
        try:
            best = float(s[:n])
            return best
        except ValueError:
            continue
    return best


def scan_doubles(line, n, prev):
    """sscanf("%lf %lf ...") semantics: values are read in order with a strtod
    prefix parse and reading stops at the first field that does not parse
    completely; fields not reached keep their previous values (the reference
    reuses one cc[] array, CMS_BasisSet.c:617-621)."""
    out = list(prev)
    toks = line.split()
    for k in range(min(n, len(toks))):
        t = toks[k]
        try:
            out[k] = float(t)
            continue
        except ValueError:
            pass
        m = re.match(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?", t)
        if m:
            out[k] = float(m.group(0))
        break
    return out


def parse_xyz(path):
    with open(path) as f:
        lines = f.readlines()
    if not lines:
        raise InputError("empty xyz")
    try:
        natoms = int(lines[0].split()[0])
    except (IndexError, ValueError):
        raise InputError("bad atom count line")
    if natoms <= 0:
        raise InputError("bad atom count")
    if len(lines) < 2:
        raise InputError("missing comment line")
    charge = int(c_atof(lines[1]))
    syms, coords = [], []
    seen = False
    for line in lines[2:]:
        tok = line.split()
        if not tok:
            if seen:
                raise InputError("blank line after atoms")
            continue
        if not tok[0][0].isalpha():
            continue
        seen = True
        if len(tok[0]) > 2 or tok[0] not in ETABLE:
            raise InputError("atom %s is not supported" % tok[0])
        if len(tok) < 4:
            raise InputError("short atom line")
        syms.append(tok[0])
        coords.append([float(tok[1]) * A2BOHR, float(tok[2]) * A2BOHR, float(tok[3]) * A2BOHR])
    if len(syms) != natoms:
        raise InputError("natoms %d does not match the header %d" % (len(syms), natoms))
    coords = np.array(coords, dtype=np.float64).reshape(natoms, 3)
    Z = np.array([ETABLE.index(s) + 1 for s in syms], dtype=np.float64)
    enuc = 0.0
    for a in range(natoms):
        for b in range(a + 1, natoms):
            enuc += Z[a] * Z[b] / math.sqrt(float(np.sum((coords[a] - coords[b]) ** 2)))
    return dict(natoms=natoms, symbols=syms, coords=coords, Z=Z, charge=charge,
                nelectrons=int(Z.sum()), enuc=enuc)


def parse_gbs(path):
    with open(path) as f:
        lines = f.readlines()
    first = lines[0].split()
    btype = first[0] if first else ""
    cartesian = btype != "spherical"
    elements = {}
    cc_state = [0.0, 0.0, 0.0, 0.0]
    i = 1
    n = len(lines)
    while i < n:
        line = lines[i]
        i += 1
        if not line[:1].isalpha():
            continue
        el = line.split()[0]
        if el not in ETABLE:
            raise InputError("atom %s in basis is not supported" % el)
        shells = []
        while i < n:
            line = lines[i]
            i += 1
            if line[:1] == "*":
                break
            if not line[:1].isalpha():
                continue
            tok = line.split()
            letters, nexp = tok[0], int(tok[1])
            if nexp <= 0 or len(letters) > 3:
                raise InputError("invalid shell line")
            prims = []
            for _ in range(nexp):
                pl = lines[i]
                i += 1
                if pl[:1] == "*" or pl[:1].isalpha():
                    raise InputError("wrong gbs format")
                cc_state = scan_doubles(pl, 4, cc_state)
                prims.append(list(cc_state))
            for k, ch in enumerate(letters):
                if ch not in MTABLE:
                    raise InputError("shell %s not supported" % letters)
                am = MTABLE.index(ch)
                exps = np.array([p[0] for p in prims])
                cc = np.array([p[1 + k] for p in prims])
                shells.append((am, exps, cc))
        elements[el] = shells
    return dict(cartesian=cartesian, elements=elements, path=path)


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def normalise_shell(am, exps, cc):
    """x^l-normalised contracted coefficients (primitive norms folded in)."""
    exps = np.asarray(exps, dtype=np.float64)
    cc = np.asarray(cc, dtype=np.float64)
    df = dfact(2 * am - 1)
    prim = (2.0 * exps / math.pi) ** 0.75 * (4.0 * exps) ** (am / 2.0) / math.sqrt(df)
    c = cc * prim
    p = exps[:, None] + exps[None, :]
    ov = df / (2.0 * p) ** am * (math.pi / p) ** 1.5
    norm = float(c @ ov @ c)
    return c / math.sqrt(norm)


def ncart(l):
    return (l + 1) * (l + 2) // 2


def load_sad(gbs, el):
    shells = gbs["elements"][el]
    nf = sum(ncart(s[0]) for s in shells) if gbs["cartesian"] else sum(2 * s[0] + 1 for s in shells)
    fname = os.path.join(os.path.dirname(gbs["path"]), el + ".dat")
    g = np.zeros((nf, nf))
    if not os.path.exists(fname):
        return g, False
    with open(fname) as f:
        vals = []
        for _ in range(nf * nf):
            line = f.readline()
            if not line:
                return np.zeros((nf, nf)), False
            vals.append(c_atof(line))
    g = np.array(vals).reshape(nf, nf)
    if np.any(g - g.T > 1e-12):
        return np.zeros((nf, nf)), False
    return g, True


class System:
    """A molecule in a basis set and an auxiliary basis set."""

    def __init__(self, gbs_path, aux_path, xyz_path):
        self.mol = parse_xyz(xyz_path)
        self.gbs = parse_gbs(gbs_path)
        self.aux = parse_gbs(aux_path)
        self.basis = self._shells(self.gbs)
        self.auxbasis = self._shells(self.aux)
        self.nbf = int(self.basis["bf_start"][-1])
        self.naux = int(self.auxbasis["bf_start"][-1])
        self.n_occ = (self.mol["nelectrons"] - self.mol["charge"]) // 2
        self.sad = self._sad()

    def _shells(self, gbs):
        am, nprim, off, xyz, alpha, coef, atom = [], [], [], [], [], [], []
        bf_start = [0]
        for a, sym in enumerate(self.mol["symbols"]):
            if sym not in gbs["elements"]:
                raise InputError("element %s missing from basis" % sym)
            for (l, exps, cc) in gbs["elements"][sym]:
                am.append(l)
                nprim.append(len(exps))
                off.append(len(alpha))
                xyz.append(self.mol["coords"][a])
                alpha.extend(exps.tolist())
                coef.extend(normalise_shell(l, exps, cc).tolist())
                atom.append(a)
                bf_start.append(bf_start[-1] + ncart(l))
        return dict(am=np.array(am, dtype=np.int32), nprim=np.array(nprim, dtype=np.int32),
                    prim_off=np.array(off, dtype=np.int32),
                    xyz=np.array(xyz, dtype=np.float64).reshape(-1, 3),
                    alpha=np.array(alpha, dtype=np.float64), coef=np.array(coef, dtype=np.float64),
                    atom=np.array(atom, dtype=np.int32), bf_start=np.array(bf_start, dtype=np.int32),
                    nshells=len(am))

    def _sad(self):
        D = np.zeros((self.nbf, self.nbf))
        b = self.basis
        for a, sym in enumerate(self.mol["symbols"]):
            idx = np.nonzero(b["atom"] == a)[0]
            s0 = int(b["bf_start"][idx[0]])
            g, _ = load_sad(self.gbs, sym)
            n = g.shape[0]
            D[s0:s0 + n, s0:s0 + n] = g
        return D
