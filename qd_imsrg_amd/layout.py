"""Host view of the blocked operator layout, read from the library's own
channel tables (qd_layout_table), plus the synthetic inputs and the
reference flop count the benchmark is quoted on.

Layout (the reference's ``vectorize`` order, src/ME_general.f90:1871-1989):
E0, fhh(n,n), fpp(m',m'), fph(m',n), then per (M_L, M_S) block the six
sub-blocks hhhh, pppp, pphh, phph, ppph, phhh, all column-major.
"""
from __future__ import annotations

import numpy as np

SUB_HHHH, SUB_PPPP, SUB_PPHH, SUB_PHPH, SUB_PPPH, SUB_PHHH = range(6)


def _table(nbody: int, nshell: int, which: int) -> np.ndarray:
    from . import _lib
    L = _lib.load()
    n = int(L.qd_layout_table(int(nbody), int(nshell), which, None, 0))
    _lib.check(min(n, 0))
    out = np.zeros(max(n, 1), dtype=np.int64)
    _lib.check(int(L.qd_layout_table(int(nbody), int(nshell), which, out.ctypes.data_as(_lib.i64p), n)) - n)
    return out[:n]


class Layout:
    """Sizes and offsets of the channel structure of N electrons in R shells
    (host tables only: no device is touched)."""

    def __init__(self, nbody: int, nshell: int):
        t = lambda which: _table(nbody, nshell, which)
        dims = t(0).reshape(-1, 3)
        self.nh, self.np_, self.nb = (dims[:, k].copy() for k in range(3))
        self.off = t(1).reshape(-1, 6)
        cc = t(4).reshape(-1, 2)
        self.cc_r, self.cc_nb = cc[:, 0].copy(), cc[:, 1].copy()
        self.neq, self.cc_size, self.nblock, self.ncc = (int(x) for x in t(6))
        sp = t(7).reshape(-1, 2)
        self.ml, self.ms = sp[:, 0].copy(), sp[:, 1].copy()
        self.M = len(self.ml)
        self.n = int(nbody)
        self.mp = self.M - self.n
        self.off_fhh = 1
        self.off_fpp = 1 + self.n * self.n
        self.off_fph = self.off_fpp + self.mp * self.mp

    def n2b(self) -> int:
        """Two-body elements: neq minus E0 and the three one-body arrays."""
        return self.neq - self.off_fph - self.mp * self.n

    def gamma_od_norm(self, v: np.ndarray) -> float:
        """Frobenius norm of the off-diagonal part f_ph and every Vpphh block."""
        f = v[self.off_fph:self.off_fph + self.mp * self.n]
        s = float(np.dot(f, f))
        for q in range(self.nblock):
            k, m = int(self.off[q, SUB_PPHH]), int(self.np_[q]) * int(self.nh[q])
            s += float(np.dot(v[k:k + m], v[k:k + m]))
        return s ** 0.5

    def reference_gemm_flops(self) -> dict:
        """2*m*n*k summed over the reference's dgemm calls for one
        commutator_222 at this channel structure: matrix_mult222
        (src/commutators.f90:999-1127; ladder = pp/hh products, skinny = the
        K = nh class, mixed = products through ph pairs) and construct_cc_ints
        (src/ME_general.f90:948-949; two r x r x nb products per CC block)."""
        lad = skinny = mixed = 0
        for h, p, b in zip(self.nh.tolist(), self.np_.tolist(), self.nb.tolist()):
            if h * p:
                lad += 2 * (h * h * p + p * p * h + 2 * p * h * p)
                skinny += 2 * (2 * p * h * h)
            if h:
                skinny += 2 * h ** 3
            if p:
                lad += 2 * p ** 3
            if b * p * h:
                mixed += 2 * (2 * b * h * p + 2 * p * b * h)
            if b * h:
                skinny += 2 * (2 * b * h * h)
                mixed += 2 * b * b * h
            if b * p:
                mixed += 2 * (2 * p * b * p + b * b * p)
        pandya = sum(4 * r * r * nb for r, nb in zip(self.cc_r.tolist(), self.cc_nb.tolist()))
        return dict(ladder=lad, skinny=skinny, mixed=mixed, pandya=pandya, total=lad + skinny + mixed + pandya)


def random_operator(lay: Layout, rng: np.random.Generator, herm: int) -> np.ndarray:
    """Seeded synthetic operator: one-body entries N(0,1) on (m_l, m_s)-allowed
    positions, two-body block entries N(0,1)/sqrt(block dim); fhh, fpp and the
    diagonal sub-blocks (hhhh, pppp, phph) (anti)symmetric to ``herm``."""
    n, mp = lay.n, lay.mp
    v = np.zeros(lay.neq)
    v[0] = rng.standard_normal()
    allow = (lay.ml[:, None] == lay.ml[None, :]) & (lay.ms[:, None] == lay.ms[None, :])

    def one(rows, cols, r0, c0, sym):
        a = rng.standard_normal((rows, cols)) * allow[r0:r0 + rows, c0:c0 + cols]
        return 0.5 * (a + herm * a.T) if sym else a

    fhh, fpp, fph = one(n, n, 0, 0, True), one(mp, mp, n, n, True), one(mp, n, n, 0, False)
    v[lay.off_fhh:lay.off_fhh + n * n] = fhh.T.reshape(-1)
    v[lay.off_fpp:lay.off_fpp + mp * mp] = fpp.T.reshape(-1)
    v[lay.off_fph:lay.off_fph + mp * n] = fph.T.reshape(-1)
    for q in range(lay.nblock):
        nh, np_, nb = int(lay.nh[q]), int(lay.np_[q]), int(lay.nb[q])
        nt = nh + np_ + nb
        if nt == 0:
            continue
        o = lay.off[q]
        for k, (r, c, sym) in enumerate([(nh, nh, True), (np_, np_, True), (np_, nh, False),
                                         (nb, nb, True), (np_, nb, False), (nb, nh, False)]):
            if r * c == 0:
                continue
            a = rng.standard_normal((r, c)) * (1.0 / np.sqrt(nt))
            if sym:
                a = 0.5 * (a + herm * a.T)
            v[o[k]:o[k] + r * c] = a.T.reshape(-1)
    return v
