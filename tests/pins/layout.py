"""Flat ``vectorize``-order operator <-> dense (E0, f, G) tensors.

The unpacking follows the reference's element accessors: f_elem
(src/ME_general.f90:1780-1813) for one-body entries -- f_hp = herm * fph^T --
and v_elem (src/ME_general.f90:1557-1680) for two-body entries: the missing
sub-blocks are herm times the transpose of the stored ones, and G is
antisymmetric in each pair.  Used only to compare the blocked oracle with the
layout-free textbook commutator.
"""
from __future__ import annotations

import numpy as np


def _mat(v, off, rows, cols):
    return v[off:off + rows * cols].reshape(cols, rows).T  # column-major


def flat_to_dense(ch, v, herm):
    n, M, mp = ch.n, ch.M, ch.mp
    f = np.zeros((M, M))
    f[:n, :n] = _mat(v, ch.off_fhh, n, n)
    f[n:, n:] = _mat(v, ch.off_fpp, mp, mp)
    fph = _mat(v, ch.off_fph, mp, n)
    f[n:, :n] = fph
    f[:n, n:] = herm * fph.T
    G = np.zeros((M, M, M, M))
    for q, b in enumerate(ch.blocks):
        nh, np_, nb = int(ch.nh[q]), int(ch.np_[q]), int(ch.nb[q])
        pairs = np.concatenate([b["qnhh"], b["qnph"], b["qnpp"]]).reshape(-1, 2)
        nt = len(pairs)
        if nt == 0:
            continue
        o = ch.off[q]
        B = np.zeros((nt, nt))
        h, p = slice(0, nh), slice(nh + nb, nt)
        x = slice(nh, nh + nb)
        Vhhhh = _mat(v, o[0], nh, nh); Vpppp = _mat(v, o[1], np_, np_)
        Vpphh = _mat(v, o[2], np_, nh); Vphph = _mat(v, o[3], nb, nb)
        Vppph = _mat(v, o[4], np_, nb); Vphhh = _mat(v, o[5], nb, nh)
        B[h, h] = Vhhhh; B[x, h] = Vphhh; B[p, h] = Vpphh
        B[h, x] = herm * Vphhh.T; B[x, x] = Vphph; B[p, x] = Vppph
        B[h, p] = herm * Vpphh.T; B[x, p] = herm * Vppph.T; B[p, p] = Vpppp
        i, j = pairs[:, 0][:, None], pairs[:, 1][:, None]
        k, l = pairs[:, 0][None, :], pairs[:, 1][None, :]
        G[i, j, k, l] = B
        G[j, i, k, l] = -B
        G[i, j, l, k] = -B
        G[j, i, l, k] = B
    return float(v[0]), f, G


def dense_to_flat(ch, E0, f, G):
    n, mp = ch.n, ch.mp
    v = np.zeros(ch.neq)
    v[0] = E0
    put = lambda off, A: v.__setitem__(slice(off, off + A.size), A.T.reshape(-1))
    put(ch.off_fhh, f[:n, :n])
    put(ch.off_fpp, f[n:, n:])
    put(ch.off_fph, f[n:, :n])
    for q, b in enumerate(ch.blocks):
        o = ch.off[q]
        HH, PH, PP = b["qnhh"], b["qnph"], b["qnpp"]
        sub = lambda L, R: G[L[:, 0][:, None], L[:, 1][:, None], R[:, 0][None, :], R[:, 1][None, :]]
        if len(HH):
            put(o[0], sub(HH, HH))
        if len(PP):
            put(o[1], sub(PP, PP))
        if len(PP) and len(HH):
            put(o[2], sub(PP, HH))
        if len(PH):
            put(o[3], sub(PH, PH))
        if len(PP) and len(PH):
            put(o[4], sub(PP, PH))
        if len(PH) and len(HH):
            put(o[5], sub(PH, HH))
    return v


def random_flat(ch, rng, herm, dense_f=False, scale=None):
    """Random operator in the blocked layout with the symmetry the reference
    assumes: fhh, fpp and the diagonal sub-blocks (hhhh, pppp, phph) are
    (anti)symmetric to ``herm``.  Two-body entries ~ N(0,1)/sqrt(ntot) per
    block; one-body entries N(0,1) on (m_l, m_s)-allowed positions (or dense)."""
    n, mp = ch.n, ch.mp
    v = np.zeros(ch.neq)
    v[0] = rng.standard_normal()
    ml, ms = ch.ml, ch.ms
    allow = (ml[:, None] == ml[None, :]) & (ms[:, None] == ms[None, :])
    def one(rows, cols, r0, c0, sym):
        A = rng.standard_normal((rows, cols))
        if not dense_f:
            A = A * allow[r0:r0 + rows, c0:c0 + cols]
        if sym:
            A = 0.5 * (A + herm * A.T)
        return A
    fhh = one(n, n, 0, 0, True); fpp = one(mp, mp, n, n, True); fph = one(mp, n, n, 0, False)
    v[ch.off_fhh:ch.off_fhh + n * n] = fhh.T.reshape(-1)
    v[ch.off_fpp:ch.off_fpp + mp * mp] = fpp.T.reshape(-1)
    v[ch.off_fph:ch.off_fph + mp * n] = fph.T.reshape(-1)
    for q in range(ch.nblock):
        nh, np_, nb = int(ch.nh[q]), int(ch.np_[q]), int(ch.nb[q])
        nt = nh + np_ + nb
        if nt == 0:
            continue
        s = (1.0 / np.sqrt(nt)) if scale is None else scale
        o = ch.off[q]
        for k, (r, c, sym) in enumerate([(nh, nh, True), (np_, np_, True), (np_, nh, False),
                                         (nb, nb, True), (np_, nb, False), (nb, nh, False)]):
            if r * c == 0:
                continue
            A = rng.standard_normal((r, c)) * s
            if sym:
                A = 0.5 * (A + herm * A.T)
            v[o[k]:o[k] + r * c] = A.T.reshape(-1)
    return v
