"""Arbitrary-precision Obara-Saika integrals (test helper).

An independent formulation (Obara & Saika, J. Chem. Phys. 84, 3963 (1986);
Head-Gordon & Pople, J. Chem. Phys. 89, 5777 (1988)) of the integrals the
oracle computes with the McMurchie-Davidson scheme: vertical recursion on the
bra and ket, horizontal recursion to the second centre, Boys function from
the lower incomplete gamma function.  Used only to pin the oracle on small
shells.
"""
import functools

import mpmath as mp

mp.mp.dps = 32


def boys(m, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    a = m + mp.mpf(1) / 2
    return mp.gammainc(a, 0, T) / (2 * T ** a)


def cart_list(l):
    out = []
    for i in range(l, -1, -1):
        for j in range(l - i, -1, -1):
            out.append((i, j, l - i - j))
    return out


def _dec(v, i):
    return tuple(x - (1 if k == i else 0) for k, x in enumerate(v))


def _inc(v, i):
    return tuple(x + (1 if k == i else 0) for k, x in enumerate(v))


def _first_nonzero(v):
    for k, x in enumerate(v):
        if x > 0:
            return k
    return -1


def overlap_prim(a, b, A, B, alpha, beta):
    p = alpha + beta
    P = [(alpha * A[k] + beta * B[k]) / p for k in range(3)]
    res = mp.mpf(1)
    for k in range(3):
        XPA, XPB = P[k] - A[k], P[k] - B[k]
        X = A[k] - B[k]

        @functools.lru_cache(None)
        def S(i, j):
            if i < 0 or j < 0:
                return mp.mpf(0)
            if i == 0 and j == 0:
                return mp.sqrt(mp.pi / p) * mp.e ** (-alpha * beta / p * X * X)
            if i > 0:
                return XPA * S(i - 1, j) + ((i - 1) * S(i - 2, j) + j * S(i - 1, j - 1)) / (2 * p)
            return XPB * S(i, j - 1) + (i * S(i - 1, j - 1) + (j - 1) * S(i, j - 2)) / (2 * p)
        res *= S(a[k], b[k])
    return res


def kinetic_prim(a, b, A, B, alpha, beta):
    """T = 1/2 <grad a | grad b>: each derivative of x^i exp(-alpha x^2) is
    i x^(i-1) - 2 alpha x^(i+1), giving four shifted 1D overlaps per axis."""
    p = alpha + beta

    def s1d(k, i, j):
        X = A[k] - B[k]
        P = (alpha * A[k] + beta * B[k]) / p
        XPA, XPB = P - A[k], P - B[k]

        @functools.lru_cache(None)
        def S(ii, jj):
            if ii < 0 or jj < 0:
                return mp.mpf(0)
            if ii == 0 and jj == 0:
                return mp.sqrt(mp.pi / p) * mp.e ** (-alpha * beta / p * X * X)
            if ii > 0:
                return XPA * S(ii - 1, jj) + ((ii - 1) * S(ii - 2, jj) + jj * S(ii - 1, jj - 1)) / (2 * p)
            return XPB * S(ii, jj - 1) + (ii * S(ii - 1, jj - 1) + (jj - 1) * S(ii, jj - 2)) / (2 * p)
        return S(i, j)

    def t1d(k, i, j):
        return (i * j * s1d(k, i - 1, j - 1) - 2 * alpha * j * s1d(k, i + 1, j - 1)
                - 2 * beta * i * s1d(k, i - 1, j + 1) + 4 * alpha * beta * s1d(k, i + 1, j + 1)) / 2

    sx, sy, sz = (s1d(k, a[k], b[k]) for k in range(3))
    tx, ty, tz = (t1d(k, a[k], b[k]) for k in range(3))
    return tx * sy * sz + sx * ty * sz + sx * sy * tz


def nuclear_prim(a, b, A, B, alpha, beta, C):
    p = alpha + beta
    P = [(alpha * A[k] + beta * B[k]) / p for k in range(3)]
    AB = [A[k] - B[k] for k in range(3)]
    PA = [P[k] - A[k] for k in range(3)]
    PC = [P[k] - C[k] for k in range(3)]
    K = mp.e ** (-alpha * beta / p * sum(x * x for x in AB))
    T = p * sum(x * x for x in PC)

    @functools.lru_cache(None)
    def V(aa, bb, m):
        if min(aa) < 0 or min(bb) < 0:
            return mp.mpf(0)
        i = _first_nonzero(bb)
        if i >= 0:
            return V(_inc(aa, i), _dec(bb, i), m) + AB[i] * V(aa, _dec(bb, i), m)
        i = _first_nonzero(aa)
        if i < 0:
            return 2 * mp.pi / p * K * boys(m, T)
        a1 = _dec(aa, i)
        r = PA[i] * V(a1, bb, m) - PC[i] * V(a1, bb, m + 1)
        if a1[i] > 0:
            a2 = _dec(a1, i)
            r += a1[i] / (2 * p) * (V(a2, bb, m) - V(a2, bb, m + 1))
        return r
    return V(tuple(a), tuple(b), 0)


def eri_prim(a, b, c, d, A, B, C, D, al, be, ga, de):
    p, q = al + be, ga + de
    P = [(al * A[k] + be * B[k]) / p for k in range(3)]
    Q = [(ga * C[k] + de * D[k]) / q for k in range(3)]
    W = [(p * P[k] + q * Q[k]) / (p + q) for k in range(3)]
    rho = p * q / (p + q)
    AB = [A[k] - B[k] for k in range(3)]
    CD = [C[k] - D[k] for k in range(3)]
    PA = [P[k] - A[k] for k in range(3)]
    QC = [Q[k] - C[k] for k in range(3)]
    WP = [W[k] - P[k] for k in range(3)]
    WQ = [W[k] - Q[k] for k in range(3)]
    KAB = mp.e ** (-al * be / p * sum(x * x for x in AB))
    KCD = mp.e ** (-ga * de / q * sum(x * x for x in CD))
    T = rho * sum((P[k] - Q[k]) ** 2 for k in range(3))
    base = 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * KAB * KCD

    @functools.lru_cache(None)
    def vrr(aa, cc, m):
        if min(aa) < 0 or min(cc) < 0:
            return mp.mpf(0)
        i = _first_nonzero(aa)
        if i >= 0:
            a1 = _dec(aa, i)
            r = PA[i] * vrr(a1, cc, m) + WP[i] * vrr(a1, cc, m + 1)
            if a1[i] > 0:
                a2 = _dec(a1, i)
                r += a1[i] / (2 * p) * (vrr(a2, cc, m) - rho / p * vrr(a2, cc, m + 1))
            if cc[i] > 0:
                r += cc[i] / (2 * (p + q)) * vrr(a1, _dec(cc, i), m + 1)
            return r
        i = _first_nonzero(cc)
        if i < 0:
            return base * boys(m, T)
        c1 = _dec(cc, i)
        r = QC[i] * vrr(aa, c1, m) + WQ[i] * vrr(aa, c1, m + 1)
        if c1[i] > 0:
            c2 = _dec(c1, i)
            r += c1[i] / (2 * q) * (vrr(aa, c2, m) - rho / q * vrr(aa, c2, m + 1))
        return r

    @functools.lru_cache(None)
    def hrr(aa, bb, cc, dd):
        if min(aa) < 0 or min(bb) < 0 or min(cc) < 0 or min(dd) < 0:
            return mp.mpf(0)
        i = _first_nonzero(bb)
        if i >= 0:
            return hrr(_inc(aa, i), _dec(bb, i), cc, dd) + AB[i] * hrr(aa, _dec(bb, i), cc, dd)
        i = _first_nonzero(dd)
        if i >= 0:
            return hrr(aa, bb, _inc(cc, i), _dec(dd, i)) + CD[i] * hrr(aa, bb, cc, _dec(dd, i))
        return vrr(aa, cc, 0)
    return hrr(tuple(a), tuple(b), tuple(c), tuple(d))


# ---- contracted shells: shell = (am, centre, exps, coefs) ----
def _contract(fn, shells, comps):
    import itertools
    tot = mp.mpf(0)
    for prims in itertools.product(*[range(len(s[2])) for s in shells]):
        cf = mp.mpf(1)
        for s, k in zip(shells, prims):
            cf *= mp.mpf(s[3][k])
        tot += cf * fn(comps, [mp.mpf(s[2][k]) for s, k in zip(shells, prims)])
    return tot


def _mpv(x):
    return [mp.mpf(float(v)) for v in x]


def overlap(sa, sb):
    A, B = _mpv(sa[1]), _mpv(sb[1])
    return [[_contract(lambda c, e: overlap_prim(c[0], c[1], A, B, e[0], e[1]), [sa, sb], (a, b))
             for b in cart_list(sb[0])] for a in cart_list(sa[0])]


def kinetic(sa, sb):
    A, B = _mpv(sa[1]), _mpv(sb[1])
    return [[_contract(lambda c, e: kinetic_prim(c[0], c[1], A, B, e[0], e[1]), [sa, sb], (a, b))
             for b in cart_list(sb[0])] for a in cart_list(sa[0])]


def nuclear(sa, sb, Zs, centres):
    A, B = _mpv(sa[1]), _mpv(sb[1])
    out = []
    for a in cart_list(sa[0]):
        row = []
        for b in cart_list(sb[0]):
            v = mp.mpf(0)
            for Z, Cc in zip(Zs, centres):
                Cm = _mpv(Cc)
                v -= mp.mpf(float(Z)) * _contract(lambda c, e: nuclear_prim(c[0], c[1], A, B, e[0], e[1], Cm),
                                                  [sa, sb], (a, b))
            row.append(v)
        out.append(row)
    return out


def eri3(sa, sb, sc):
    """(ab|c) with the ket paired with an exponent-0 s function."""
    A, B, C = _mpv(sa[1]), _mpv(sb[1]), _mpv(sc[1])
    z = (0, 0, 0)
    out = []
    for a in cart_list(sa[0]):
        for b in cart_list(sb[0]):
            for c in cart_list(sc[0]):
                out.append(_contract(lambda cm, e: eri_prim(cm[0], cm[1], cm[2], z, A, B, C, C, e[0], e[1], e[2], 0),
                                     [sa, sb, sc], (a, b, c)))
    return out


def eri2(sc, sd):
    C, D = _mpv(sc[1]), _mpv(sd[1])
    z = (0, 0, 0)
    return [_contract(lambda cm, e: eri_prim(cm[0], z, cm[1], z, C, C, D, D, e[0], 0, e[1], 0), [sc, sd], (c, d))
            for c in cart_list(sc[0]) for d in cart_list(sd[0])]


def eri4(sa, sb, sc, sd, ia, ib, ic, id_):
    A, B, C, D = _mpv(sa[1]), _mpv(sb[1]), _mpv(sc[1]), _mpv(sd[1])
    comps = (cart_list(sa[0])[ia], cart_list(sb[0])[ib], cart_list(sc[0])[ic], cart_list(sd[0])[id_])
    return _contract(lambda cm, e: eri_prim(cm[0], cm[1], cm[2], cm[3], A, B, C, D, e[0], e[1], e[2], e[3]),
                     [sa, sb, sc, sd], comps)
