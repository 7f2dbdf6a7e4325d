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


def magnus_flow(o: Oracle, h0: np.ndarray, stp: float = R4_0P7, conv: float = R4_1EM9,
                max_iter: int = 100000, omega0: np.ndarray | None = None):
    """Returns dict(hs, omega, s, e0, iterations, rejected, bch_terms, trace)."""
    if omega0 is None:
        g = o.zeros()
        hs = h0.copy()
        terms = 0
        crit = abs(o.mbpt2(h0))
    else:
        g = omega0.copy()
        hs, terms = o.bch_expand(g, -1, h0)
        crit = abs(o.mbpt2(hs))
    eta = o.build_white(hs, 1)
    nrm1 = o.norm(eta)
    s = 0.0
    it = rej = steps = 0
    trace = []
    while crit > conv and steps < max_iter:
        steps += 1
        g0 = g.copy()
        dg = eta.copy()                       # MAGNUS_EXPAND as shipped
        g = o.add(g, 1.0, dg, stp)            # euler_step
        s += stp
        hs0 = hs.copy()
        hs, nt = o.bch_expand(g, -1, h0)
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
        crit = abs(o.mbpt2(hs))
        it += 1
    return dict(hs=hs, omega=g, s=s, e0=hs[0], iterations=it, rejected=rej, bch_terms=terms,
                crit=crit, trace=np.array(trace).reshape(-1, 6))
