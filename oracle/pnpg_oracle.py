"""FP64 restatement of the trial-by-trial PNPG alpha step -- TEST INFRASTRUCTURE ONLY.

npg/PNPG.m:50-167 with forcePositive and u = 0 (prox = projection on alpha >= 0), the
majorization test of utils/Utils.m:470-480, stepShrnk 0.5, stepIncre 0.9, gamma 2,
a 1/4, adaptive steps and restarts as in the reference class defaults; one call of
``trial`` is one pass of the line-search loop (a failed test shrinks the step and keeps
the iterate; a restart resets theta).  ``func(x) -> (f, grad)``.
"""
from __future__ import annotations

import numpy as np


class PnpgOracle:
    def __init__(self, func, alpha0, t0):
        self.func = func
        self.alpha = np.array(alpha0, np.float64)
        self.pre = self.alpha.copy()
        self.t = float(t0)
        self.preT = float(t0)
        self.theta = 0.0
        self.cost = np.inf
        self.p = 0
        self.ppp = 0
        self.cumu = 0
        self.cumuTol = 4
        self.incStep = False
        self.log = []

    def trial(self):
        if self.ppp == 0:
            self.incStep = False
            if self.cumu >= self.cumuTol:
                self.t *= 0.9
                self.cumu = 0
                self.incStep = True
        self.ppp += 1
        B = self.t / self.preT
        newTheta = 1.0 if self.p == 0 else 0.5 + np.sqrt(0.25 + B * self.theta ** 2)
        coef = (self.theta - 1.0) / newTheta
        xbar = np.maximum(self.alpha + coef * (self.alpha - self.pre), 0.0)
        f_old, g = self.func(xbar)
        newx = np.maximum(xbar - g / self.t, 0.0)
        f_new, _ = self.func(newx, grad=False)
        d = newx - xbar
        holds = (f_new - f_old) <= d @ g + 0.5 * self.t * (d @ d)
        if not holds and self.ppp <= 20 and self.t > 0:
            self.t /= 0.5
            if self.incStep:
                self.cumuTol += 4
                self.incStep = False
            self.log.append(("reject", f_new))
            return
        if holds and self.p > 0 and f_new > self.cost and coef != 0.0:
            self.theta = 1.0
            self.ppp = 0
            self.log.append(("restart", f_new))
            return
        self.cumu = self.cumu + 1 if self.ppp == 1 else 0
        self.theta = newTheta
        self.cost = f_new
        self.preT = self.t
        self.p += 1
        self.ppp = 0
        self.pre = self.alpha
        self.alpha = newx
        self.log.append(("accept", f_new))
