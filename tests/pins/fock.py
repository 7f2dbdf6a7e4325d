"""Fock-space brute force: exact commutators of normal-ordered operators.

Pins ``textbook.commutator``.  Operators are built as dense matrices on the
full 2^M Fock space from their normal-ordered 0/1/2-body parts (normal order
with respect to the reference state with holes 0..n-1 filled).  The exact
[A, B] = AB - BA contains a 3-body part; its 0/1/2-body normal-ordered
components -- the IM-SRG(2) truncation -- are read off matrix elements between
states whose quasi-particle counts add to at most 4, where no normal-ordered
3-body operator (6 quasi-particle operators) can contribute.
"""
from __future__ import annotations

import itertools

import numpy as np


class Fock:
    def __init__(self, M: int, nholes: int):
        self.M, self.n = M, nholes
        self.dim = 1 << M
        self.states = np.arange(self.dim, dtype=np.int64)
        self.pop = np.array([bin(s).count("1") for s in range(self.dim)], dtype=np.int64)
        self.phi = (1 << nholes) - 1

    def _apply(self, idx, dag, st, amp):
        bit = (st >> idx) & 1
        ok = (bit == 0) if dag else (bit == 1)
        sign = 1 - 2 * (self.pop[st & ((1 << idx) - 1)] & 1)
        new = (st | (1 << idx)) if dag else (st & ~(1 << idx))
        return new[ok], (amp * sign)[ok]

    def no_string(self, ops):
        """Dense matrix of the normal-ordered product of ``ops`` = [(idx, dag)]
        written left to right (quasi-particle creators moved left, sign kept)."""
        is_cr = [(d and i >= self.n) or ((not d) and i < self.n) for (i, d) in ops]
        order = [k for k in range(len(ops)) if is_cr[k]] + [k for k in range(len(ops)) if not is_cr[k]]
        inv = sum(1 for a in range(len(order)) for b in range(a + 1, len(order)) if order[a] > order[b])
        sign = -1.0 if inv & 1 else 1.0
        st, amp = self.states.copy(), np.full(self.dim, sign)
        src = self.states.copy()
        for k in reversed(order):
            i, d = ops[k]
            bit = (st >> i) & 1
            ok = (bit == 0) if d else (bit == 1)
            sg = 1 - 2 * (self.pop[st & ((1 << i) - 1)] & 1)
            st = ((st | (1 << i)) if d else (st & ~(1 << i)))[ok]
            amp = (amp * sg)[ok]
            src = src[ok]
        Mx = np.zeros((self.dim, self.dim))
        np.add.at(Mx, (st, src), amp)
        return Mx

    def build(self, op):
        """Fock matrix of the normal-ordered operator op = (E0, f, G)."""
        E0, f, G = op
        M = self.M
        X = E0 * np.eye(self.dim)
        for p in range(M):
            for q in range(M):
                if f[p, q] != 0.0:
                    X += f[p, q] * self.no_string([(p, True), (q, False)])
        for p, q in itertools.combinations(range(M), 2):
            for r, s in itertools.combinations(range(M), 2):
                if G[p, q, r, s] != 0.0:
                    X += G[p, q, r, s] * self.no_string([(p, True), (q, True), (s, False), (r, False)])
        return X

    def _signature(self, ops):
        ket = self.phi
        for (i, d) in ops:
            if (not d) and i >= self.n:
                ket |= 1 << i
            if d and i < self.n:
                ket &= ~(1 << i)
        E = self.no_string(ops)
        col = E[:, ket]
        bra = int(np.flatnonzero(col)[0])
        return bra, ket, col[bra], E

    def extract(self, X):
        """Normal-ordered (E0, f, G) of the Fock matrix X, up to 2-body."""
        M = self.M
        E0 = X[self.phi, self.phi]
        X1 = X - E0 * np.eye(self.dim)
        f = np.zeros((M, M))
        Efs = {}
        for p in range(M):
            for q in range(M):
                bra, ket, amp, E = self._signature([(p, True), (q, False)])
                f[p, q] = X1[bra, ket] / amp
                Efs[(p, q)] = E
        X2 = X1 - sum(f[p, q] * Efs[(p, q)] for (p, q) in Efs)
        G = np.zeros((M, M, M, M))
        for p, q in itertools.combinations(range(M), 2):
            for r, s in itertools.combinations(range(M), 2):
                bra, ket, amp, _ = self._signature([(p, True), (q, True), (s, False), (r, False)])
                g = X2[bra, ket] / amp
                G[p, q, r, s] = g
                G[q, p, r, s] = -g
                G[p, q, s, r] = -g
                G[q, p, s, r] = g
        return E0, f, G
