"""TDA spectrum and EOM-IM-SRG restated for the oracle.  TEST INFRASTRUCTURE
ONLY (see oracle/oracle.py).

tda_matrix: build_TDAmat + calc_TDA (src/operators.f90:7-124): the (hole i,
particle a) pairs with a below the cut (1-based a = N+1 .. cutshell) and
m_l(a) - m_l(i) = ML, m_s(a) - m_s(i) = MS, hole outer; matrix element
d_ij f_ab - d_ab f_ji + <aj|V|ib>.  Eigenvalues as dsyev('V','U') gives them.

eom_matrix: the dense EOM matrix of transfer (ML, MS) built column by column
from the C oracle's matvec_prod (imsrg_oracle.c, o_eom_matvec); eom_states:
its lowest eigenvalues (LANCZOS_DIAGONALIZE asks ARPACK for the smallest
algebraic ones, src/EOM_IMSRG.f90:232-277) and the ph weight of each
eigenvector (sum(fph**2), src/EOM_IMSRG.f90:44-46).
"""
from __future__ import annotations

import numpy as np

from .oracle import Oracle


def tda_labels(ch, ml: int, ms: int, cut: int):
    return [(i, a) for i in range(ch.n) for a in range(ch.n, min(cut, ch.M))
            if ch.ml[a] - ch.ml[i] == ml and ch.ms[a] - ch.ms[i] == ms]


def tda_matrix(o: Oracle, v: np.ndarray, herm: int, ml: int, ms: int, cut: int) -> np.ndarray:
    lab = tda_labels(o.ch, ml, ms, cut)
    d = len(lab)
    A = np.zeros((d, d))
    for I, (i, a) in enumerate(lab):
        for J, (j, b) in enumerate(lab):
            x = o.velem(v, herm, a, j, i, b)
            if i == j:
                x += o.felem(v, herm, a, b)
            if a == b:
                x -= o.felem(v, herm, j, i)
            A[I, J] = x
    return A


def tda_spectrum(o: Oracle, v, herm, ml, ms, cut) -> np.ndarray:
    A = tda_matrix(o, v, herm, ml, ms, cut)
    if A.size == 0:
        return np.zeros(0)
    U = np.triu(A)
    return np.linalg.eigvalsh(U + np.triu(A, 1).T)


def eom_matrix(o: Oracle, h: np.ndarray, herm: int, ml: int, ms: int) -> np.ndarray:
    n = o.eom_dim(ml, ms)
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = o.eom_matvec(h, herm, ml, ms, e)
    return A


def eom_states(o: Oracle, h, herm, ml, ms, nev):
    A = eom_matrix(o, h, herm, ml, ms)
    sps = n_ph(o.ch, ml, ms)
    w, U = np.linalg.eigh(0.5 * (A + A.T))
    return w[:nev], (U[:sps, :nev] ** 2).sum(axis=0), A


def n_ph(ch, ml, ms) -> int:
    return sum(1 for a in range(ch.n, ch.M) for i in range(ch.n)
               if ch.ml[a] - ch.ml[i] == ml and ch.ms[a] - ch.ms[i] == ms)
