"""Magnus flow driver of the oracle -- restates ``s_evolve``
(src/magnus_IMSRG.f90:194-248) with ``MAGNUS_EXPAND`` as shipped (dOmega =
eta, :342-345), ``euler_step`` (:372-384) and ``BCH_EXPAND`` (:253-318, in
imsrg_oracle.c).  TEST INFRASTRUCTURE ONLY (see oracle/oracle.py).
"""
from __future__ import annotations

import numpy as np

from .oracle import Oracle

# REAL(4) literals of the reference widened to REAL(8)
R4_0P7 = 0.699999988079071        # stp = .7            src/magnus_IMSRG.f90:124
R4_1EM9 = 9.999999717180685e-10   # conv_criteria = 1e-9 src/magnus_IMSRG.f90:10
R4_1EM3 = 0.0010000000474974513   # stp/2 + 1e-3         src/magnus_IMSRG.f90:226


BCH_COF = (1.0, 0.5, 0.166666666666666666, 0.04166666666666666, 0.0083333333333333333,
           .001388888888888, 1.984126984e-4, 2.48015873e-5, 2.755731922e-6, 2.755731922e-7,
           2.505210839e-8)          # src/magnus_IMSRG.f90:265-267
R4_1EM5 = 9.999999747378752e-06   # conv = 1e-5 (REAL(4)), src/magnus_IMSRG.f90:259


def bch_expand_pandya(o: Oracle, g: np.ndarray, h: np.ndarray):
    """BCH_EXPAND with every nested commutator's CC inputs filled by calc_cc
    (the Pandya term the shipped Magnus driver leaves out)."""
    hs = h.copy()
    ad = h.copy()
    for k in range(2, 13):
        int1 = hs
        nxt = o.commutator(g, -1, ad, 1, pandya=True)
        hs = o.add(int1, 1.0, nxt, BCH_COF[k - 2])
        if o.norm(nxt) * BCH_COF[k - 2] / o.norm(int1) < R4_1EM5:
            return hs, k - 1
        ad = nxt
    return hs, 11


DIIS_START = 1e-4


def diis_weights(B: np.ndarray) -> np.ndarray | None:
    """Pulay weights: minimise |sum c_i r_i|^2 with sum c_i = 1 (bordered
    system of the Gram matrix B, scaled by its largest diagonal)."""
    n = B.shape[0]
    A = np.zeros((n + 1, n + 1))
    A[:n, :n] = B / B.diagonal().max()
    A[:n, n] = A[n, :n] = 1.0
    rhs = np.zeros(n + 1)
    rhs[n] = 1.0
    try:
        c = np.linalg.solve(A, rhs)[:n]
    except np.linalg.LinAlgError:
        return None
    return c if np.all(np.isfinite(c)) else None


def magnus_flow(o: Oracle, h0: np.ndarray, stp: float = R4_0P7, conv: float = R4_1EM9,
                max_iter: int = 100000, omega0: np.ndarray | None = None, pandya: bool = False,
                diis: int = 0, stop_rule: str = "mbpt2", conv_deds: float = 1e-8):
    """Returns dict(hs, omega, s, e0, iterations, rejected, bch_terms, trace).
    diis > 0: once |MBPT2| <= DIIS_START, Pulay acceleration over the last
    ``diis`` (<= 6) iterates of the map Omega -> Omega + stp eta(Omega)
    (history dropped when ||eta|| grows past twice its smallest value since);
    the reference's step with its rejection rule otherwise (the library's
    qd_magnus_flow with qd_flow_params.diis; not in the reference).
    stop_rule "deds": stop when |dE/ds| <= conv_deds instead, dE/ds being the
    zero-body part of [eta(s), H(s)] (commutator_110 + commutator_220 of the
    White generator and H(s)); the library's qd_flow_params.stop_rule = 1."""
    def deds(eta_, hs_):
        return o.commutator_110(eta_, -1, hs_, 1) + o.commutator_220(eta_, -1, hs_, 1)

    def crit_of(eta_, hs_):
        return abs(deds(eta_, hs_)) if stop_rule == "deds" else abs(o.mbpt2(hs_))
    bch = (lambda g, h: bch_expand_pandya(o, g, h)) if pandya else (lambda g, h: o.bch_expand(g, -1, h))
    if stop_rule == "deds":
        conv = conv_deds
    if omega0 is None:
        g = o.zeros()
        hs = h0.copy()
        terms = 0
    else:
        g = omega0.copy()
        hs, terms = bch(g, h0)
    # step-rejection scale: from H(0) also for a warm start (the library's
    # qd_magnus_flow does the same; the reference always starts cold)
    nrm1 = o.norm(o.build_white(h0, 1))
    eta = o.build_white(hs, 1)
    crit = crit_of(eta, hs)
    s = 0.0
    it = rej = steps = 0
    trace = []
    m = min(max(diis, 0), 6)
    HO, HR, B, count, best = [None] * max(m, 1), [None] * max(m, 1), np.zeros((max(m, 1),) * 2), 0, 0.0
    while crit > conv and steps < max_iter:
        steps += 1
        if m > 0 and crit <= DIIS_START:
            slot = count % m
            HO[slot], HR[slot] = g.copy(), eta.copy()
            count += 1
            n = min(count, m)
            for i in range(n):
                B[slot, i] = B[i, slot] = float(np.dot(HR[slot][1:], HR[i][1:]))
            c = diis_weights(B[:n, :n]) if n >= 2 else None
            if c is None:
                g = o.add(g, 1.0, eta, stp)
            else:
                g = sum(ci * HO[i] + (ci * stp) * HR[i] for i, ci in enumerate(c))
            s += stp
            hs, nt = bch(g, h0)
            terms += nt
            eta = o.build_white(hs, 1)
            nrm2 = o.norm(eta)
            trace.append((it, s, hs[0], nrm1, nrm2, crit))
            if count == 1 or nrm2 < best:
                best = nrm2
            if nrm2 > 2.0 * best:
                count = 0
            crit = crit_of(eta, hs)
            it += 1
            continue
        count = 0
        g0 = g.copy()
        dg = eta.copy()                       # MAGNUS_EXPAND as shipped
        g = o.add(g, 1.0, dg, stp)            # euler_step
        s += stp
        hs0 = hs.copy()
        hs, nt = bch(g, h0)
        terms += nt
        eta0 = eta.copy()
        eta = o.build_white(hs, 1, gen=eta.copy())
        nrm2 = o.norm(eta)
        trace.append((it, s, hs[0], nrm1, nrm2, crit))
        if nrm1 * 1.5 < nrm2:
            s -= stp
            g, hs, eta = g0, hs0, eta0
            stp = stp / 2.0 + R4_1EM3
            rej += 1
            continue
        crit = crit_of(eta, hs)
        it += 1
    return dict(hs=hs, omega=g, s=s, e0=hs[0], iterations=it, rejected=rej, bch_terms=terms,
                crit=crit, mbpt2=o.mbpt2(hs), deds=deds(eta, hs), trace=np.array(trace).reshape(-1, 6))
