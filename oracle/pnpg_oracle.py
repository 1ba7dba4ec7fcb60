"""FP64 restatement of the trial-by-trial PNPG alpha step -- TEST INFRASTRUCTURE ONLY.

npg/PNPG.m:50-167 with forcePositive and u = 0 (prox = projection on alpha >= 0), the
majorization test of utils/Utils.m:470-480, stepShrnk 0.5, stepIncre 0.9, gamma 2,
a 1/4, adaptive steps and restarts as in the reference class defaults; one call of
``trial`` is one pass of the line-search loop (a failed test shrinks the step and keeps
the iterate; a restart resets theta).  ``func(x) -> (f, grad)``.  ``PnpgWaveletOracle``
adds the 'wvltADMM' prox (npg/admm.m via oracle/wavelet_oracle.py) with the admmTol /
maxInnerItr schedule of PNPG.m:113-146.
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


class PnpgWaveletOracle(PnpgOracle):
    """npg/PNPG.m:50-167 with the 'wvltADMM' prox (bhc/beamhardenSpline.m:204-209):
    newX = admm(xbar - grad/t, u/t, admmTol*difAlpha, maxInnerItr, alpha) on the images
    of ``shape``, objective f + u ||mdwt(x)||_1, and the admmTol / maxInnerItr schedule of
    PNPG.m:113-146.  A failed line search sets theta = 1 (reset() without re-estimating
    t).  ``h``, ``L``: the wavelet filter and levels."""

    def __init__(self, func, alpha0, t0, u, h, L, shape):
        super().__init__(func, alpha0, t0)
        from oracle import wavelet_oracle as W
        self.W = W
        self.u, self.h, self.L, self.shape = float(u), h, L, shape
        self.admmTol = 1e-2
        self.maxInnerItr = 100
        self.maxPossibleInnerItr = 1000
        self.difAlpha = 1.0
        self.innerSearch = 0

    def _pen(self, x):
        return self.W.l1_penalty(x.reshape(self.shape), self.h, self.L)

    def _prox(self, a, u, tol):
        return self.W.admm(a, u, self.h, self.L, tol, self.maxInnerItr,
                           self.alpha.reshape(self.shape))

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
        a = (xbar - g / self.t).reshape(self.shape)
        newx, self.innerSearch = self._prox(a, self.u / self.t, self.admmTol * self.difAlpha)
        newx = newx.ravel()
        f_new, _ = self.func(newx, grad=False)
        d = newx - xbar
        holds = (f_new - f_old) <= d @ g + 0.5 * self.t * (d @ d)
        goodMM = True
        if not holds:
            if self.ppp <= 20 and self.t > 0:
                self.t /= 0.5
                if self.incStep:
                    self.cumuTol += 4
                    self.incStep = False
                self.log.append(("reject", f_new))
                return
            goodMM = False
        newObj = f_new + self.u * self._pen(newx)
        objBar = f_old + self.u * self._pen(xbar)
        if newObj - self.cost > 0:
            if goodMM and np.abs(xbar - self.alpha).sum() != 0:
                self.theta = 1.0
                self.ppp = 0
                self.log.append(("restart", newObj))
                return
            if (not goodMM) or objBar < newObj:
                if not goodMM:
                    self.theta = 1.0
                if self.innerSearch < self.maxInnerItr and self.admmTol > 1e-6:
                    self.admmTol /= 10
                    self.ppp = 0
                    self.log.append(("admmTol", newObj))
                    return
                if self.innerSearch >= self.maxInnerItr and \
                        self.maxInnerItr < self.maxPossibleInnerItr:
                    self.maxInnerItr *= 10
                    self.ppp = 0
                    self.log.append(("maxInnerItr", newObj))
                    return
                newObj = self.cost
                newx = self.alpha
        self.theta = newTheta
        self.cost = newObj
        nb = np.sqrt(newx @ newx)
        self.difAlpha = np.sqrt((self.alpha - newx) @ (self.alpha - newx)) / max(nb, 1e-10)
        self.cumu = self.cumu + 1 if self.ppp == 1 else 0
        self.preT = self.t
        self.p += 1
        self.ppp = 0
        self.pre = self.alpha
        self.alpha = newx
        self.log.append(("accept", newObj))


class PnpgTVOracle(PnpgWaveletOracle):
    """The same trials with the TV prox ('tvl1' / 'tviso': utils/TV.m denoise with
    innerThresh = admmTol*difAlpha, penalty utils/tlv.m)."""

    def __init__(self, func, alpha0, t0, u, tv, shape):
        super().__init__(func, alpha0, t0, u, None, 0, shape)
        from oracle import tv_oracle as T
        self.T = T
        self.tv = tv

    def _pen(self, x):
        return self.T.tlv(x.reshape(self.shape), self.tv)

    def _prox(self, a, u, tol):
        return self.T.denoise(a, u, self.tv, tol, self.maxInnerItr)

