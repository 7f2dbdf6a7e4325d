"""Textbook IM-SRG(2) commutator on dense tensors -- an independent pin.

Normal-ordered operators O = E0 + sum f_pq {p+ q} + 1/4 sum G_pqrs {p+ q+ s r},
G antisymmetric in (p,q) and in (r,s); occupations n_a = 1 for holes.  The
0/1/2-body parts of [A, B] in the IM-SRG(2) truncation are the fundamental
commutators of Hergert et al., Phys. Rep. 621, 165 (2016), written here with
einsum on full M^4 tensors.  Nothing here knows the reference's block layout;
``fock.py`` checks these formulas against exact Fock-space commutators.
"""
from __future__ import annotations

import numpy as np


def _es(spec, *ops):
    return np.einsum(spec, *ops, optimize=True)


def commutator(A, B, nocc, pandya: bool = True):
    """[A, B] for A = (E0, f, G); returns (C0, C1, C2)."""
    _, fa, Ga = A
    _, fb, Gb = B
    n = np.asarray(nocc, dtype=float)
    nb = 1.0 - n
    # ---- zero body
    dn = n[:, None] - n[None, :]                      # n_a - n_b
    C0 = _es("ab,ab,ba->", dn, fa, fb)
    w4 = _es("a,b,c,d->abcd", n, n, nb, nb)
    C0 += 0.25 * (_es("abcd,abcd,cdab->", w4, Ga, Gb) - _es("abcd,abcd,cdab->", w4, Gb, Ga))
    # ---- one body
    C1 = fa @ fb - fb @ fa
    C1 += _es("ab,ab,biaj->ij", dn, fa, Gb) - _es("ab,ab,biaj->ij", dn, fb, Ga)
    w3 = _es("a,b,c->abc", n, n, nb) + _es("a,b,c->abc", nb, nb, n)
    C1 += 0.5 * (_es("abc,ciab,abcj->ij", w3, Ga, Gb) - _es("abc,ciab,abcj->ij", w3, Gb, Ga))
    # ---- two body: 1b-2b
    T = _es("ia,ajkl->ijkl", fa, Gb) - _es("ia,ajkl->ijkl", fb, Ga)
    C2 = T - T.transpose(1, 0, 2, 3)
    U = _es("ak,ijal->ijkl", fa, Gb) - _es("ak,ijal->ijkl", fb, Ga)
    C2 -= U - U.transpose(0, 1, 3, 2)
    # ---- pp/hh ladders: 1/2 sum_ab (1 - n_a - n_b)(A_ijab B_abkl - B_ijab A_abkl)
    M = len(n)
    lad = (1.0 - n[:, None] - n[None, :]).reshape(1, M * M)
    a2, b2 = Ga.reshape(M * M, M * M), Gb.reshape(M * M, M * M)
    C2 += 0.5 * ((a2 * lad) @ b2 - (b2 * lad) @ a2).reshape(M, M, M, M)
    # ---- particle-hole (Pandya) term: -sum_ab (n_a - n_b)(1 - P_ij)(1 - P_kl) A_bjal B_aibk
    if pandya:
        X = Ga.transpose(1, 3, 0, 2).reshape(M * M, M * M)          # [(j,l),(b,a)]
        Y = (Gb.transpose(2, 0, 1, 3) * dn.T[:, :, None, None]).reshape(M * M, M * M)  # [(b,a),(i,k)]
        P = (X @ Y).reshape(M, M, M, M).transpose(2, 0, 3, 1)        # [i,j,k,l]
        C2 -= P - P.transpose(1, 0, 2, 3) - P.transpose(0, 1, 3, 2) + P.transpose(1, 0, 3, 2)
    return C0, C1, C2


def random_operator(M, rng, herm=None, scale=1.0):
    """Random normal-ordered operator with antisymmetric G; if ``herm`` is
    +1/-1 the operator is made (anti)hermitian."""
    f = rng.standard_normal((M, M)) * scale
    G = rng.standard_normal((M, M, M, M)) * scale
    G = G - G.transpose(1, 0, 2, 3)
    G = G - G.transpose(0, 1, 3, 2)
    if herm is not None:
        f = 0.5 * (f + herm * f.T)
        G = 0.5 * (G + herm * G.transpose(2, 3, 0, 1))
    return (float(rng.standard_normal()), f, G)
