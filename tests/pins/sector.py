"""A normal-ordered operator as a matrix on the determinants with K particles.

With occupations n_h = 1 for the holes, the normal-ordered operator
E0 + sum f_pq {p+ q} + 1/4 sum G_pqrs {p+ q+ s r} equals, in vacuum order,
  c + sum f'_pq p+ q + 1/4 sum G_pqrs p+ q+ s r,
  c = E0 - sum_h f_hh + 1/2 sum_hh' G_hh'hh',   f'_pq = f_pq - sum_h G_phqh
(Wick's theorem).  Determinants are bit masks; signs count the occupied
orbitals below the one acted on.  Cheap for the few-particle sectors of the
EOM pins (C(12, 3) = 220 determinants) where the 2^M Fock matrices of
fock.py are too large.
"""
from __future__ import annotations

import itertools

import numpy as np


def _annihilate(mask, k):
    if not (mask >> k) & 1:
        return None, 0
    sign = -1 if bin(mask & ((1 << k) - 1)).count("1") & 1 else 1
    return mask & ~(1 << k), sign


def _create(mask, k):
    if (mask >> k) & 1:
        return None, 0
    sign = -1 if bin(mask & ((1 << k) - 1)).count("1") & 1 else 1
    return mask | (1 << k), sign


def sector_matrix(op, nholes: int, K: int):
    """(determinants, matrix) of op = (E0, f, G) on the K-particle sector."""
    E0, f, G = op
    M = f.shape[0]
    h = range(nholes)
    c = E0 - sum(f[i, i] for i in h) + 0.5 * sum(G[i, j, i, j] for i in h for j in h)
    fp = f - np.einsum("phqh->pq", G[:, :nholes, :, :nholes])
    dets = [sum(1 << k for k in occ) for occ in itertools.combinations(range(M), K)]
    index = {d: i for i, d in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))
    for col, d in enumerate(dets):
        H[col, col] += c
        occ = [k for k in range(M) if (d >> k) & 1]
        for q in occ:
            m1, s1 = _annihilate(d, q)
            for p in range(M):
                if fp[p, q] == 0.0:
                    continue
                m2, s2 = _create(m1, p)
                if m2 is not None:
                    H[index[m2], col] += fp[p, q] * s1 * s2
        for r, s in itertools.combinations(occ, 2):      # p+ q+ s r with r < s, p < q
            m1, s1 = _annihilate(d, r)
            m2, s2 = _annihilate(m1, s)
            for p, q in itertools.combinations(range(M), 2):
                g = G[p, q, r, s]
                if g == 0.0:
                    continue
                m3, s3 = _create(m2, q)
                if m3 is None:
                    continue
                m4, s4 = _create(m3, p)
                if m4 is not None:
                    H[index[m4], col] += g * s1 * s2 * s3 * s4
    return dets, H
