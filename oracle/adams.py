"""CPU oracle of the ODE path (test infrastructure only).

The reference integrates the IM-SRG flow dH/ds = [eta(H), H] with the
Shampine-Gordon variable-order, variable-step Adams PECE code (modified
divided differences, local extrapolation, order 1..12) in
src/adams_ode.f90 (ode :7, de :218, step :562, intrp :1187), driven by
src/gs_IMSRG.f90:100-120 (RHS dGam :223-273).  This module restates that
published algorithm (Shampine and Gordon, "Computer Solution of Ordinary
Differential Equations: the Initial Value Problem", 1975) in numpy for a
generic right-hand side f(t, y) -> y'.

Pinned by closed-form problems (tests/test_adams_oracle.py): exponential
decay, a harmonic oscillator (rotation, energy conserved), y' = -2 t y and a
linear system with known matrix exponential, each to the requested
tolerance, plus the step/order bookkeeping the algorithm guarantees (stops
exactly at tout, interpolates between steps, raises the order on smooth
problems, flags too many steps).
"""
from __future__ import annotations

import math

import numpy as np

GSTR = (0.5, 0.0833, 0.0417, 0.0264, 0.0188, 0.0143, 0.0114, 0.00936, 0.00789, 0.00679, 0.00592, 0.00524, 0.00468)
TWO = tuple(2.0 ** (i + 1) for i in range(13))
U = np.finfo(np.float64).eps
MAXNUM = 500


class Adams:
    """State of one integration (the reference's work/iwork arrays)."""

    def __init__(self, f, neqn: int):
        self.f, self.n = f, neqn
        self.phi = np.zeros((16, neqn))          # phi[i-1] is the reference's phi(:, i)
        self.yy = np.zeros(neqn)
        self.p = np.zeros(neqn)
        self.yp = np.zeros(neqn)
        self.wt = np.zeros(neqn)
        self.psi = np.zeros(12)
        self.alpha = np.zeros(12)
        self.beta = np.zeros(12)
        self.sig = np.zeros(13)
        self.v = np.zeros(12)
        self.w = np.zeros(12)
        self.g = np.zeros(13)
        self.x = self.h = self.hold = self.told = self.delsgn = 0.0
        self.k = self.kold = self.ns = 0
        self.isnold = 0
        self.start = self.phase1 = self.nornd = False
        self.nfe = 0          # right-hand side evaluations
        self.nsteps = 0       # successful steps
        self.log = []         # (x, h, k) per successful step

    def _f(self, x, y):
        self.nfe += 1
        return np.asarray(self.f(x, y), dtype=np.float64)

    # ---- STEP (src/adams_ode.f90:562-1185): one step of size h from x ----
    def step(self, eps):
        fouru, twou = 4.0 * U, 2.0 * U
        phi, y, wt = self.phi, self.yy, self.wt
        if abs(self.h) < fouru * abs(self.x):
            self.h = math.copysign(fouru * abs(self.x), self.h)
            return True, eps
        p5eps = 0.5 * eps
        rnd = twou * math.sqrt(float(np.sum((y / wt) ** 2)))
        if p5eps < rnd:
            return True, 2.0 * rnd * (1.0 + fouru)
        g, sig = self.g, self.sig
        g[0], g[1], sig[0] = 1.0, 0.5, 1.0
        if self.start:
            self.yp = self._f(self.x, y)
            phi[0] = self.yp
            phi[1] = 0.0
            total = math.sqrt(float(np.sum((self.yp / wt) ** 2)))
            absh = abs(self.h)
            if eps < 16.0 * total * self.h * self.h:
                absh = 0.25 * math.sqrt(eps / total)
            self.h = math.copysign(max(absh, fouru * abs(self.x)), self.h)
            self.hold, self.k, self.kold = 0.0, 1, 0
            self.start, self.phase1, self.nornd = False, True, True
            if p5eps <= 100.0 * rnd:
                self.nornd = False
                phi[14] = 0.0
        ifail = 0
        psi, alpha, beta, v, w = self.psi, self.alpha, self.beta, self.v, self.w
        while True:
            # block 1: coefficients of the formulas for this step
            k = self.k
            kp1, kp2, km1, km2 = k + 1, k + 2, k - 1, k - 2
            h = self.h
            if h != self.hold:
                self.ns = 0
            if self.ns <= self.kold:
                self.ns += 1
            ns = self.ns
            nsp1 = ns + 1
            if k >= ns:
                beta[ns - 1] = 1.0
                alpha[ns - 1] = 1.0 / ns
                temp1 = h * ns
                sig[nsp1 - 1] = 1.0
                for i in range(nsp1, k + 1):
                    temp2 = psi[i - 2]
                    psi[i - 2] = temp1
                    beta[i - 1] = beta[i - 2] * psi[i - 2] / temp2
                    temp1 = temp2 + h
                    alpha[i - 1] = h / temp1
                    sig[i] = i * alpha[i - 1] * sig[i - 1]
                psi[k - 1] = temp1
                if ns <= 1:
                    for iq in range(1, k + 1):
                        v[iq - 1] = 1.0 / (iq * (iq + 1))
                        w[iq - 1] = v[iq - 1]
                else:
                    if k > self.kold:
                        v[k - 1] = 1.0 / (k * kp1)
                        for j in range(1, ns - 1):
                            i = k - j
                            v[i - 1] -= alpha[j] * v[i]
                    temp5 = alpha[ns - 1]
                    for iq in range(1, kp1 - ns + 1):
                        v[iq - 1] -= temp5 * v[iq]
                        w[iq - 1] = v[iq - 1]
                    g[nsp1 - 1] = w[0]
                for i in range(ns + 2, kp1 + 1):
                    temp6 = alpha[i - 2]
                    for iq in range(1, kp2 - i + 1):
                        w[iq - 1] -= temp6 * w[iq]
                    g[i - 1] = w[0]
            # block 2: predict, evaluate, estimate the local errors
            for i in range(nsp1, k + 1):
                phi[i - 1] *= beta[i - 1]
            phi[kp2 - 1] = phi[kp1 - 1]
            phi[kp1 - 1] = 0.0
            p = np.zeros(self.n)
            for j in range(1, k + 1):
                i = kp1 - j
                p += g[i - 1] * phi[i - 1]
                phi[i - 1] += phi[i]
            if self.nornd:
                p = y + h * p
            else:
                tau = h * p - phi[14]
                p = y + tau
                phi[15] = (p - y) - tau
            self.p = p
            xold = self.x
            self.x = self.x + h
            absh = abs(h)
            self.yp = self._f(self.x, p)
            yp = self.yp
            temp4 = yp - phi[0]
            erkm2 = erkm1 = 0.0
            if km2 > 0:
                erkm2 = absh * sig[km1 - 1] * GSTR[km2 - 1] * math.sqrt(float(np.sum(((phi[km1 - 1] + temp4) / wt) ** 2)))
            if km2 >= 0:
                erkm1 = absh * sig[k - 1] * GSTR[km1 - 1] * math.sqrt(float(np.sum(((phi[k - 1] + temp4) / wt) ** 2)))
            serk = math.sqrt(float(np.sum((temp4 / wt) ** 2)))
            err = absh * serk * (g[k - 1] - g[kp1 - 1])
            erk = absh * serk * sig[kp1 - 1] * GSTR[k - 1]
            knew = k
            if km2 > 0:
                if max(erkm1, erkm2) <= erk:
                    knew = km1
            elif km2 == 0:
                if erkm1 <= 0.5 * erk:
                    knew = km1
            if err <= eps:
                break
            # block 3: unsuccessful step; restore and retry smaller
            self.phase1 = False
            self.x = xold
            for i in range(1, k + 1):
                phi[i - 1] = (1.0 / beta[i - 1]) * (phi[i - 1] - phi[i])
            for i in range(2, k + 1):
                psi[i - 2] = psi[i - 1] - h
            ifail += 1
            temp2 = 0.5
            if ifail > 3 and p5eps < 0.25 * erk:
                temp2 = math.sqrt(p5eps / erk)
            if ifail >= 3:
                knew = 1
            self.h = temp2 * h
            self.k = knew
            if abs(self.h) < fouru * abs(self.x):
                self.h = math.copysign(fouru * abs(self.x), self.h)
                return True, eps + eps
        # block 4: successful step; correct, evaluate, update differences,
        # choose order and step size for the next step
        self.kold, self.hold = k, h
        temp1 = h * g[kp1 - 1]
        if self.nornd:
            y[:] = p + temp1 * (yp - phi[0])
        else:
            rho = temp1 * (yp - phi[0]) - phi[15]
            y[:] = p + rho
            phi[14] = (y - p) - rho
        self.yp = self._f(self.x, y)
        yp = self.yp
        phi[kp1 - 1] = yp - phi[0]
        phi[kp2 - 1] = phi[kp1 - 1] - phi[kp2 - 1]
        for i in range(1, k + 1):
            phi[i - 1] += phi[kp1 - 1]
        erkp1 = 0.0
        if knew == km1 or k == 12:
            self.phase1 = False
        action = None
        if self.phase1:
            action = "raise"
        elif knew == km1:
            action = "lower"
        elif kp1 > self.ns:
            action = "keep"
        else:
            erkp1 = absh * GSTR[kp1 - 1] * math.sqrt(float(np.sum((phi[kp2 - 1] / wt) ** 2)))
            if k > 1:
                if erkm1 <= min(erk, erkp1):
                    action = "lower"
                elif erkp1 >= erk or k == 12:
                    action = "keep"
                else:
                    action = "raise"
            else:
                action = "keep" if erkp1 >= 0.5 * erk else "raise"
        if action == "raise":
            self.k, erk = kp1, erkp1
        elif action == "lower":
            self.k, erk = km1, erkm1
        hnew = h + h
        if not self.phase1 and p5eps < erk * TWO[self.k]:
            hnew = h
            if p5eps < erk:
                r = (p5eps / erk) ** (1.0 / (self.k + 1))
                hnew = absh * max(0.5, min(0.9, r))
                hnew = math.copysign(max(hnew, fouru * abs(self.x)), h)
        self.h = hnew
        self.nsteps += 1
        self.log.append((self.x, h, k))
        return False, eps

    # ---- INTRP (src/adams_ode.f90:1187-1312) ----
    def intrp(self, xout):
        hi = xout - self.x
        ki = self.kold + 1
        kip1 = ki + 1
        g = np.zeros(13)
        rho = np.zeros(13)
        w = np.zeros(13)
        g[0] = rho[0] = 1.0
        for i in range(1, ki + 1):
            w[i - 1] = 1.0 / i
        term = 0.0
        for j in range(2, ki + 1):
            psijm1 = self.psi[j - 2]
            gamma = (hi + term) / psijm1
            eta = hi / psijm1
            for i in range(1, kip1 - j + 1):
                w[i - 1] = gamma * w[i - 1] - eta * w[i]
            g[j - 1] = w[0]
            rho[j - 1] = gamma * rho[j - 2]
            term = psijm1
        yout = np.zeros(self.n)
        ypout = np.zeros(self.n)
        for j in range(1, ki + 1):
            i = kip1 - j
            yout += g[i - 1] * self.phi[i - 1]
            ypout += rho[i - 1] * self.phi[i - 1]
        return self.yy + hi * yout, ypout

    # ---- DE / ODE (src/adams_ode.f90:7-560): integrate from t to tout ----
    def ode(self, y, t, tout, relerr, abserr, iflag):
        """Returns (y, t, relerr, abserr, iflag) as the reference's ode."""
        fouru = 4.0 * U
        if self.n < 1 or t == tout or relerr < 0 or abserr < 0:
            return y, t, relerr, abserr, 6
        eps = max(relerr, abserr)
        if eps <= 0 or iflag == 0:
            return y, t, relerr, abserr, 6
        isn = 1 if iflag > 0 else -1
        iflag = abs(iflag)
        if iflag != 1 and (t != self.told or iflag < 2 or iflag > 5):
            return y, t, relerr, abserr, 6
        dl = tout - t
        absdel = abs(dl)
        tend = tout if isn < 0 else t + 10.0 * dl
        nostep = kle4 = 0
        stiff = False
        releps, abseps = relerr / eps, abserr / eps
        if iflag == 1 or self.isnold < 0 or self.delsgn * dl <= 0.0:
            self.start = True
            self.x = t
            self.yy = np.array(y, dtype=np.float64, copy=True)
            self.delsgn = math.copysign(1.0, dl)
            self.h = math.copysign(max(abs(tout - self.x), fouru * abs(self.x)), tout - self.x)
        while True:
            if abs(self.x - t) >= absdel:
                yout, _ = self.intrp(tout)
                self.told, self.isnold = tout, isn
                return yout, tout, relerr, abserr, 2
            if isn < 0 and abs(tout - self.x) < fouru * abs(self.x):
                h = tout - self.x
                yp = self._f(self.x, self.yy)
                self.told, self.isnold = tout, isn
                return self.yy + h * yp, tout, relerr, abserr, 2
            if nostep >= MAXNUM:
                self.told, self.isnold = self.x, 1
                return self.yy.copy(), self.x, relerr, abserr, isn * (5 if stiff else 4)
            self.h = math.copysign(min(abs(self.h), abs(tend - self.x)), self.h)
            self.wt = releps * np.abs(self.yy) + abseps
            crash, eps = self.step(eps)
            if crash:
                self.told, self.isnold = self.x, 1
                return self.yy.copy(), self.x, eps * releps, eps * abseps, isn * 3
            nostep += 1
            kle4 += 1
            if self.kold > 4:
                kle4 = 0
            if kle4 >= 50:
                stiff = True


# Reference driver constants (src/gs_IMSRG.f90:13,92-98), REAL(4) literals widened
GS_REL = float(np.float32(1e-8))
GS_ABS = float(np.float32(1e-6))
GS_STP = float(np.float32(0.1))
GS_CONV = float(np.float32(1e-7))


def gs_flow(o, h0, max_out: int = 100000, relerr: float = GS_REL, abserr: float = GS_ABS):
    """gs_IMSRG's decoupling loop (src/gs_IMSRG.f90:100-120): integrate
    dH/ds = dGam(H) in output intervals of stp until |dE0/E0| over one
    interval is below 1e-7.  RHS = dGam (src/gs_IMSRG.f90:223-273): White
    generator, calc_cc of H and eta, [eta, H] with the Pandya term.  As in
    the reference, each output interval calls ode with the flag the last call
    returned (3, 4 and 5 continue the integration; 6 leaves y unchanged, so
    the next test sees no energy change and the loop ends)."""
    def dgam(_t, y):
        eta = o.build_white(y, 1)
        return o.commutator(eta, -1, y, 1, pandya=True)
    a = Adams(dgam, h0.size)
    y = h0.copy()
    s, flag, crit, out = 0.0, 1, 1.0, []
    rel, ab = relerr, abserr
    flags = []
    while crit > GS_CONV and len(out) < max_out:
        e_old = y[0]
        y, s, rel, ab, flag = a.ode(y, s, s + GS_STP, rel, ab, flag)
        crit = abs((e_old - y[0]) / e_old)
        out.append((len(out) + 1, s, y[0], crit))
        flags.append(flag)
    return dict(hs=y, e0=y[0], s=s, outputs=len(out), nfe=a.nfe, nsteps=a.nsteps, flag=flag,
                flags=flags, relerr=rel, abserr=ab, trace=np.array(out), log=a.log)
