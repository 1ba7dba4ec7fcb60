"""FP64 restatement of the trial-by-trial PNPG alpha step -- TEST INFRASTRUCTURE ONLY.

npg/PNPG.m:50-167 with forcePositive and u = 0 (prox = projection on alpha >= 0), the
majorization test of utils/Utils.m:470-480, stepShrnk 0.5, stepIncre 0.9, gamma 2,
a 1/4, adaptive steps and restarts as in the reference class defaults; one call of
``trial`` is one pass of the line-search loop (a failed test shrinks the step and keeps
the iterate; a restart resets theta).  ``func(x) -> (f, grad)``.  With ``npg=True`` the
NPG variant (npg/NPG.m:45-148): FISTA theta once per iteration, adaptive increase t/4,
and in the first ``pre_steps`` iterations the step grows while the test holds.
``PnpgWaveletOracle`` adds the 'wvltADMM' prox (npg/admm.m via oracle/wavelet_oracle.py)
with the admmTol / maxInnerItr schedule of PNPG.m:113-146, ``PnpgTVOracle`` the TV prox.
"""
from __future__ import annotations

import numpy as np


class PnpgOracle:
    def __init__(self, func, alpha0, t0, npg=False, pre_steps=2, mask=None):
        """mask: support of alpha (flattened bool), pixels outside stay zero."""
        self.mask = None if mask is None else np.asarray(mask, bool).ravel()
        self.func = func if mask is None else self._masked(func)
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
        self.npg = npg
        self.pre_steps = pre_steps
        self.outer = 0
        self.goodStep = True
        self.newTheta = 1.0
        self.coef = 0.0
        self.admmTol = 1e-2
        self.max_ls = 20
        self.log = []

    def _masked(self, func):
        def f(x, grad=True):
            v, g = func(x, grad)
            return v, (g * self.mask if g is not None else None)
        return f

    def _begin(self):
        """PNPG.m:56-77 / NPG.m:48-66: theta, the adaptive step, xbar."""
        if self.ppp == 0:
            self.incStep = False
            if self.npg:
                self.outer += 1
                self.goodStep = True
                self.newTheta = (1 + np.sqrt(1 + 4 * self.theta ** 2)) / 2
                self.coef = (self.theta - 1.0) / self.newTheta
            if self.cumu >= self.cumuTol:
                if self.npg:
                    self.t = self.t * 0.5
                    self.t = self.t * 0.5
                else:
                    self.t *= 0.9
                self.cumu = 0
                self.incStep = True
        self.ppp += 1
        if not self.npg:
            B = self.t / self.preT
            self.newTheta = 1.0 if self.p == 0 else 0.5 + np.sqrt(0.25 + B * self.theta ** 2)
            self.coef = (self.theta - 1.0) / self.newTheta
        xbar = np.maximum(self.alpha + self.coef * (self.alpha - self.pre), 0.0)
        return xbar if self.mask is None else xbar * self.mask

    def _linesearch(self, holds, f_new):
        """0: another trial follows, 1: the test holds, 2: given up (goodMM false)."""
        if holds:
            if self.npg and self.outer <= self.pre_steps and self.ppp < 18 and \
                    self.goodStep and self.t > 0:
                self.t = self.t * 0.5
                self.log.append(("grow", f_new))
                return 0
            return 1
        if self.ppp <= self.max_ls and self.t > 0:
            self.t /= 0.5
            self.goodStep = False
            if self.incStep:
                self.cumuTol += 4
                self.incStep = False
            self.log.append(("reject", f_new))
            return 0
        return 2

    def _accept(self, newx, newObj):
        self.cumu = self.cumu + 1 if self.ppp == 1 else 0
        self.theta = self.newTheta
        self.cost = newObj
        self.preT = self.t
        self.p += 1
        self.ppp = 0
        self.pre = self.alpha
        self.alpha = newx
        self.log.append(("accept", newObj))

    def trial(self):
        xbar = self._begin()
        f_old, g = self.func(xbar)
        newx = np.maximum(xbar - g / self.t, 0.0)
        f_new, _ = self.func(newx, grad=False)
        d = newx - xbar
        holds = (f_new - f_old) <= d @ g + 0.5 * self.t * (d @ d)
        ls = self._linesearch(holds, f_new)
        if ls == 0:
            return
        code = self._decide(ls, f_old, f_new, bool(np.any(xbar != self.alpha)))
        if code >= 2:
            self._accept(self.alpha if code == 3 else newx, self.cost if code == 3 else f_new)

    def _decide(self, ls, f_old, f_new, moved):
        """PNPG.m:99-160 with u = 0 after a finished line search (goodMM = ls == 1).  The
        non-negativity prox (customMax) reports one inner iteration, below maxInnerItr =
        100, so the admmTol schedule applies.  Returns 0 restart, 1 retry with admmTol / 10,
        2 take newX, 3 keep alpha ("force it to converge")."""
        good = ls == 1
        if f_new - self.cost > 0:
            if good and moved:
                self.theta = 1.0
                self.ppp = 0
                self.log.append(("restart", f_new))
                return 0
            if not good or f_old < f_new:
                if not good:
                    self.theta = 1.0
                if self.admmTol > 1e-6:
                    self.admmTol /= 10
                    self.ppp = 0
                    self.log.append(("retry", f_new))
                    return 1
                return 3
        return 2


class PnpgWaveletOracle(PnpgOracle):
    """npg/PNPG.m:50-167 with the 'wvltADMM' prox (bhc/beamhardenSpline.m:204-209):
    newX = admm(xbar - grad/t, u/t, admmTol*difAlpha, maxInnerItr, alpha) on the images
    of ``shape``, objective f + u ||mdwt(x)||_1, and the admmTol / maxInnerItr schedule of
    PNPG.m:113-146.  A given-up line search calls reset() (PNPG.m:168-172): theta = 1
    and t = min(t, hess_step(alpha)), hess_step giving stepSizeInit('hessian')
    (npg/Methods.m:108-115) at the current iterate when supplied.  ``h``, ``L``: the
    wavelet filter and levels."""

    def __init__(self, func, alpha0, t0, u, h, L, shape, npg=False, pre_steps=2, mask=None,
                 coef_mask=None, hess_step=None):
        super().__init__(func, alpha0, t0, npg, pre_steps, mask)
        self.hess_step = hess_step
        self.t_reset = 0.0
        self.resets = 0
        from oracle import wavelet_oracle as W
        self.W = W
        self.ops = None
        if mask is not None:  # Psi / Psit on the supports' vectors (ConfigCT.m:85-86)
            ck = np.ones(shape, bool) if coef_mask is None else np.asarray(coef_mask, bool)
            self.ops = W.masked_ops(h, L, np.asarray(mask, bool).reshape(shape), ck)
        self.u, self.h, self.L, self.shape = float(u), h, L, shape
        self.admmTol = 1e-2
        self.maxInnerItr = 100
        self.maxPossibleInnerItr = 1000
        self.difAlpha = 1.0
        self.innerSearch = 0

    def _pen(self, x):
        if self.ops is not None:
            return float(np.abs(self.ops[1](x[self.mask])).sum())
        return self.W.l1_penalty(x.reshape(self.shape), self.h, self.L)

    def _prox(self, a, u, tol):
        if self.ops is not None:
            v, it = self.W.admm(a.ravel()[self.mask], u, self.h, self.L, tol, self.maxInnerItr,
                                self.alpha[self.mask], ops=self.ops)
            x = np.zeros(self.mask.size)
            x[self.mask] = v
            return x.reshape(self.shape), it
        return self.W.admm(a, u, self.h, self.L, tol, self.maxInnerItr,
                           self.alpha.reshape(self.shape))

    def trial(self):
        xbar = self._begin()
        f_old, g = self.func(xbar)
        a = (xbar - g / self.t).reshape(self.shape)
        newx, self.innerSearch = self._prox(a, self.u / self.t, self.admmTol * self.difAlpha)
        newx = newx.ravel()
        f_new, _ = self.func(newx, grad=False)
        d = newx - xbar
        holds = (f_new - f_old) <= d @ g + 0.5 * self.t * (d @ d)
        ls = self._linesearch(holds, f_new)
        if ls == 0:
            return
        goodMM = ls == 1
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
                    if self.hess_step is not None:
                        th = self.hess_step(self.alpha)
                        th = 1.0 if np.isnan(th) else (np.finfo(float).eps if th == 0 else th)
                        self.t = min(self.t, th)
                        self.t_reset = th
                        self.resets += 1
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
        nb = np.sqrt(newx @ newx)
        self.difAlpha = np.sqrt((self.alpha - newx) @ (self.alpha - newx)) / max(nb, 1e-10)
        self._accept(newx, newObj)


class PnpgTVOracle(PnpgWaveletOracle):
    """The same trials with the TV prox ('tvl1' / 'tviso': utils/TV.m denoise with
    innerThresh = admmTol*difAlpha, penalty utils/tlv.m)."""

    def __init__(self, func, alpha0, t0, u, tv, shape, npg=False, pre_steps=2, hess_step=None):
        super().__init__(func, alpha0, t0, u, None, 0, shape, npg, pre_steps,
                         hess_step=hess_step)
        from oracle import tv_oracle as T
        self.T = T
        self.tv = tv

    def _pen(self, x):
        return self.T.tlv(x.reshape(self.shape), self.tv)

    def _prox(self, a, u, tol):
        return self.T.denoise(a, u, self.tv, tol, self.maxInnerItr)

