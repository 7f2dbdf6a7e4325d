"""Analytic 2D harmonic-oscillator Coulomb matrix element -- oracle side.

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py).

Restates ``v_int`` (src/ME_general.f90:2205-2276; Anisimovas & Matulis,
J. Phys.: Condens. Matter 10, 601 (1998)) with the special functions of
src/ME_general.f90:2300-2396.  Arguments are (n, m_l, m_s) triples; the value
is <i j|V|l k> in the usual notation (spin deltas s_i = s_l, s_j = s_k), in
units where the oscillator length is 1, times sqrt(hw).

``v_int`` works in binary64 with exact integer factorials, like the
reference (REAL(16) factorials, REAL(8) sums).  ``v_int_exact`` evaluates the
same sum in exact rational arithmetic times sqrt(pi/2) and is used to pin it.
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from math import comb, factorial, sqrt

import mpmath


def _gs(qi, qj, qk, ql, j1, j2, j3, j4):
    ai, aj, ak, al = abs(qi[1]), abs(qj[1]), abs(qk[1]), abs(ql[1])
    g1 = j1 + j4 + (ai + qi[1]) // 2 + (al - ql[1]) // 2
    g4 = j1 + j4 + (ai - qi[1]) // 2 + (al + ql[1]) // 2
    g2 = j2 + j3 + (aj + qj[1]) // 2 + (ak - qk[1]) // 2
    g3 = j2 + j3 + (aj - qj[1]) // 2 + (ak + qk[1]) // 2
    return g1, g2, g3, g4


def _gamma_half(twice: int):
    """Gamma(twice/2) as (rational, has_sqrt_pi)."""
    if twice % 2 == 0:
        return Fraction(factorial(twice // 2 - 1)), False
    xc = (twice - 1) // 2   # Gamma(xc + 1/2) = (2xc)! / (4^xc xc!) sqrt(pi)
    return Fraction(factorial(2 * xc), 4 ** xc * factorial(xc)), True


@lru_cache(maxsize=None)
def _inner(g1, g2, g3, g4):
    """sum over l1..l4 with l1+l2 = l3+l4 of the binomial/Gamma product, as an
    exact (rational coefficient of 1, rational coefficient of sqrt(pi))."""
    G = g1 + g2 + g3 + g4
    a = Fraction(0)
    b = Fraction(0)
    for l1 in range(g1 + 1):
        for l2 in range(g2 + 1):
            for l3 in range(g3 + 1):
                l4 = l1 + l2 - l3
                if l4 < 0 or l4 > g4:
                    continue
                lam = l1 + l2 + l3 + l4
                sgn = -1 if (g2 + g3 - l2 - l3) % 2 else 1
                c = sgn * comb(g1, l1) * comb(g2, l2) * comb(g3, l3) * comb(g4, l4)
                # Gamma(1 + lam/2) Gamma((G - lam + 1)/2); lam is even
                ga, pa = _gamma_half(2 + lam)
                gb, pb = _gamma_half(G - lam + 1)
                term = c * ga * gb
                if pa or pb:
                    b += term
                else:
                    a += term
    return a, b


def _terms(qi, qj, qk, ql):
    """Yields (G, rational prefactor, inner (a,b)) for every (j1..j4)."""
    for j1 in range(qi[0] + 1):
        for j2 in range(qj[0] + 1):
            for j3 in range(qk[0] + 1):
                for j4 in range(ql[0] + 1):
                    g = _gs(qi, qj, qk, ql, j1, j2, j3, j4)
                    G = sum(g)
                    sgn = -1 if (j1 + j2 + j3 + j4) % 2 else 1
                    pref = Fraction(sgn * comb(qi[0] + abs(qi[1]), qi[0] - j1) * comb(qj[0] + abs(qj[1]), qj[0] - j2)
                                    * comb(qk[0] + abs(qk[1]), qk[0] - j3) * comb(ql[0] + abs(ql[1]), ql[0] - j4),
                                    factorial(j1) * factorial(j2) * factorial(j3) * factorial(j4))
                    yield G, pref, _inner(*g)


def _norm2(q):
    return Fraction(factorial(q[0]), factorial(q[0] + abs(q[1])))


def v_int(qi, qj, qk, ql, hw: float = 1.0) -> float:
    """binary64 restatement of v_int (src/ME_general.f90:2205-2276)."""
    if qi[2] != ql[2] or qk[2] != qj[2]:
        return 0.0
    sm = 0.0
    sqpi = 1.77245385090551602
    for G, pref, (a, b) in _terms(qi, qj, qk, ql):
        sm_int = float(a) + float(b) * sqpi
        sm += float(pref) * 0.5 ** ((G + 1) * 0.5) * sm_int
    nrm = float(_norm2(qi) * _norm2(qj) * _norm2(qk) * _norm2(ql))
    return sqrt(hw) * sm * sqrt(nrm)


def v_int_exact(qi, qj, qk, ql, hw=1, dps: int = 50):
    """The same sum evaluated exactly (rationals, sqrt(pi), sqrt(2)) to dps digits."""
    if qi[2] != ql[2] or qk[2] != qj[2]:
        return mpmath.mpf(0)
    with mpmath.workdps(dps):
        tot = mpmath.mpf(0)
        for G, pref, (a, b) in _terms(qi, qj, qk, ql):
            inner = mpmath.mpf(a.numerator) / a.denominator + mpmath.sqrt(mpmath.pi) * (
                mpmath.mpf(b.numerator) / b.denominator)
            tot += mpmath.mpf(pref.numerator) / pref.denominator * mpmath.power(2, -mpmath.mpf(G + 1) / 2) * inner
        nrm = _norm2(qi) * _norm2(qj) * _norm2(qk) * _norm2(ql)
        return mpmath.sqrt(hw) * tot * mpmath.sqrt(mpmath.mpf(nrm.numerator) / nrm.denominator)


def spatial_table(nshell: int, hw: float = 1.0):
    """<ab|V|cd> for all spatial orbitals a,b,c,d of the first nshell shells
    with m_a + m_b = m_c + m_d (0 elsewhere): array [Nmax]*4 where the
    spatial orbital index follows nm_index (src/ME_general.f90:2140-2149):
    shell R = 2n+|m|, index R(R+1)/2 + (R+m)/2."""
    orbs = []
    for R in range(nshell):
        for m in range(-R, R + 1, 2):
            orbs.append(((R - abs(m)) // 2, m))
    N = len(orbs)
    import numpy as np
    T = np.zeros((N, N, N, N))
    for a, (na, ma) in enumerate(orbs):
        for b, (nb_, mb) in enumerate(orbs):
            for c, (nc, mc) in enumerate(orbs):
                for d, (nd, md) in enumerate(orbs):
                    if ma + mb != mc + md:
                        continue
                    # <ab|V|cd> = v_int(a, b, d, c) with equal spins
                    T[a, b, c, d] = v_int((na, ma, 1), (nb_, mb, 1), (nd, md, 1), (nc, mc, 1), hw)
    return orbs, T
