"""FP64 NumPy restatement of the beam-hardening model -- TEST INFRASTRUCTURE ONLY.

Follows, line by line:
  * ``b1_iout``  -- bhc/Spline.m:122-203 (closed form + 3rd-order Taylor branch below
                    |s| < |EPS*temp1(1)/temp1(end)|^(1/3), EPS = 1e-6)
  * ``b0_iout``  -- bhc/Spline.m:205-262
  * ``dis_iout`` -- bhc/Spline.m:264-302
  * ``poiss_l_alpha`` -- bhc/poissLAlpha.m:1-19
  * ``poiss_l_i``     -- bhc/poissLI.m:1-26
  * ``gauss_l_alpha`` -- bhc/gaussLAlpha.m:1-21
  * ``gauss_l_i``     -- bhc/gaussLI.m:1-12
  * knots / kappa grid -- bhc/beamhardenSpline.m:138-139,164-165,182
plus the moment forms ``b1_iout_stable`` / ``b0_iout_stable`` (the kernels' formulas),
the I-epsilon step ``ie_newton`` (the library's damped projected Newton, restated),
an independent minimiser ``ie_minimize`` (scipy L-BFGS-B with bounds), and the
measurement simulator ``gen_beam_harden`` (bhc/genBeamHarden.m:53-79) with
``energy_bins``.

Pinned without MATLAB (the reference's BHC code is MATLAB and cannot run here) by:
quadrature of phi_j(s) = int basis_j(kappa) exp(-kappa s) dkappa for the hat and box
bases (scipy.integrate.quad), the closed values phi_j(0), finite differences of the
dis basis' derivatives, central finite differences of the likelihoods, and the
identity g . Ie = sum(Ir - Imea) of poissLI; the simulator by its monochromatic closed
form (Ts = the largest 0.8^j with exp(-Ts kappa rho (max y - min y)) >= scale) and
hand-computed two-energy cases.  Exact MATLAB outputs of the Taylor
branches are "parity unpinned" beyond those checks.
"""
from __future__ import annotations

import numpy as np


def b1_iout(k, s, I=None, nout=3):
    """bhc/Spline.m:122-203.  I None -> basis matrices (len(s) x E)."""
    EPS = 1e-6
    k = np.asarray(k, np.float64).ravel()
    s = np.asarray(s, np.float64).ravel()
    E = len(k) - 2
    shape = (len(s),) if I is not None else (len(s), E)
    BL = np.zeros(shape)
    sBL = np.zeros(shape)
    ssBL = np.zeros(shape)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        expksC = np.exp(-k[0] * s)
        expksR = np.exp(-k[1] * s)
        s2, s3, s4, abss = s ** 2, s ** 3, s ** 4, np.abs(s)
        for i in range(E):
            a, b, c = k[i], k[i + 1], k[i + 2]
            expksL = expksC
            expksC = expksR
            expksR = np.exp(-c * s)
            temp = -(expksL - expksC) / (s2 * (a - b)) + (expksC - expksR) / (s2 * (b - c))
            t1 = np.array([
                -a / 2 + c / 2,
                (a ** 3 - b ** 3) / (6 * (a - b)) - (b ** 3 - c ** 3) / (6 * (b - c)),
                -(a ** 4 - b ** 4) / (24 * (a - b)) + (b ** 4 - c ** 4) / (24 * (b - c)),
                (a ** 5 - b ** 5) / (120 * (a - b)) - (b ** 5 - c ** 5) / (120 * (b - c)),
            ])
            thr = abs(EPS * t1[0] / t1[-1]) ** (1.0 / (len(t1) - 1))
            idx = abss < thr
            temp[idx] = t1[0] + np.stack([s[idx], s2[idx], s3[idx]], 1) @ t1[1:]
            if I is not None:
                BL += temp * I[i]
            else:
                BL[:, i] = temp
            if nout > 1:
                temp = (2 * (expksL - expksC)) / (s3 * (a - b)) + (a * expksL - b * expksC) / (s2 * (a - b)) \
                    - (2 * (expksC - expksR)) / (s3 * (b - c)) - (b * expksC - c * expksR) / (s2 * (b - c))
                t1 = np.array([
                    (a ** 3 - b ** 3) / (6 * (a - b)) - (b ** 3 - c ** 3) / (6 * (b - c)),
                    -(a ** 4 - b ** 4) / (12 * (a - b)) + (b ** 4 - c ** 4) / (12 * (b - c)),
                    (a ** 5 - b ** 5) / (40 * (a - b)) - (b ** 5 - c ** 5) / (40 * (b - c)),
                    -(a ** 6 - b ** 6) / (180 * (a - b)) + (b ** 6 - c ** 6) / (180 * (b - c)),
                ])
                thr = abs(EPS * t1[0] / t1[-1]) ** (1.0 / (len(t1) - 1))
                idx = abss < thr
                temp[idx] = t1[0] + np.stack([s[idx], s2[idx], s3[idx]], 1) @ t1[1:]
                if I is not None:
                    sBL += temp * I[i]
                else:
                    sBL[:, i] = temp
            if nout > 2:
                temp = -(6 * expksL - 6 * expksC + s2 * (a ** 2 * expksL - b ** 2 * expksC)
                         + s * (4 * a * expksL - 4 * b * expksC)) / (s4 * (a - b)) \
                    + (6 * expksC - 6 * expksR + s2 * (b ** 2 * expksC - c ** 2 * expksR)
                       + s * (4 * b * expksC - 4 * c * expksR)) / (s4 * (b - c))
                t1 = np.array([
                    -((a + b) * (a ** 2 + b ** 2)) / 12 + ((b + c) * (b ** 2 + c ** 2)) / 12,
                    (a ** 4 + a ** 3 * b + a ** 2 * b ** 2 + a * b ** 3 + b ** 4) / 20
                    + (-b ** 4 - b ** 3 * c - b ** 2 * c ** 2 - b * c ** 3 - c ** 4) / 20,
                    -((a + b) * (a ** 2 + a * b + b ** 2) * (a ** 2 - a * b + b ** 2)) / 60
                    + ((b + c) * (b ** 2 + b * c + c ** 2) * (b ** 2 - b * c + c ** 2)) / 60,
                    (a ** 6 + a ** 5 * b + a ** 4 * b ** 2 + a ** 3 * b ** 3 + a ** 2 * b ** 4 + a * b ** 5 + b ** 6) / 252
                    + (-b ** 6 - b ** 5 * c - b ** 4 * c ** 2 - b ** 3 * c ** 3 - b ** 2 * c ** 4 - b * c ** 5 - c ** 6) / 252,
                ])
                thr = abs(EPS * t1[0] / t1[-1]) ** (1.0 / (len(t1) - 1))
                idx = abss < thr
                temp[idx] = t1[0] + np.stack([s[idx], s2[idx], s3[idx]], 1) @ t1[1:]
                if I is not None:
                    ssBL += temp * I[i]
                else:
                    ssBL[:, i] = temp
    return BL, sBL, ssBL


def b1_iout_quad(k, s, j):
    """phi_j(s), phi_j'(s), phi_j''(s) by adaptive quadrature of the hat (independent pin)."""
    from scipy.integrate import quad
    a, b, c = k[j], k[j + 1], k[j + 2]

    def hat(x):
        return (x - a) / (b - a) if x <= b else (c - x) / (c - b)

    out = []
    for p in range(3):
        f = lambda x, p=p: ((-x) ** p) * hat(x) * np.exp(-x * s)  # noqa: E731
        v = quad(f, a, b, epsabs=0, epsrel=1e-13, limit=200)[0] + \
            quad(f, b, c, epsabs=0, epsrel=1e-13, limit=200)[0]
        out.append(v)
    return tuple(out)


_M3_COEF = None


def _moments(z):
    """M_n(z) = int_0^1 u^n exp(-z u) du, n = 0..3, vectorised over z (FP64).
    |z| <= 2: M_3 by its series sum_k (-z)^k / (k! (k+4)) (22 terms, truncation below
    2e-16 of M_3), then the downward recurrence M_{n-1} = (z M_n + e^-z)/n, which damps
    errors by |z|/n <= 1 per step; else the upward recurrence M_0 = (1-e^-z)/z,
    M_n = (n M_{n-1} - e^-z)/z.  Each form is evaluated only where it applies."""
    global _M3_COEF
    if _M3_COEF is None:
        from math import factorial
        _M3_COEF = [(-1.0) ** k / (factorial(k) * (k + 4)) for k in range(22)]
    z = np.asarray(z, np.float64)
    small = np.abs(z) <= 2.0
    M = np.empty((4,) + z.shape)
    zs = z[small]
    es = np.exp(-zs)
    m3 = np.full_like(zs, _M3_COEF[-1])
    for c in reversed(_M3_COEF[:-1]):
        m3 = m3 * zs + c
    m2 = (zs * m3 + es) / 3
    m1 = (zs * m2 + es) / 2
    m0 = zs * m1 + es
    zl = z[~small]
    e = np.exp(-zl)
    b0 = (1 - e) / zl
    b1 = (b0 - e) / zl
    b2 = (2 * b1 - e) / zl
    b3 = (3 * b2 - e) / zl
    for n, (lo, big) in enumerate(((m0, b0), (m1, b1), (m2, b2), (m3, b3))):
        M[n][small] = lo
        M[n][~small] = big
    return M


def b1_iout_stable(k, s, I=None, nout=3):
    """The B1 spectrum transform by segment moments, smooth in s (no branch switch):
    substituting kappa = k_j + Delta_j u in phi_j(s) = int hat_j(kappa) e^{-kappa s} dkappa.
    Same contract as b1_iout; pinned against quadrature (b1_iout_quad).  With I given,
    BL, sBL, ssBL are accumulated basis by basis (the ray x basis matrix is not formed)."""
    k = np.asarray(k, np.float64).ravel()
    s = np.asarray(s, np.float64).ravel()
    E = len(k) - 2
    if I is None:
        P = np.zeros((3, E, len(s)))
        add = lambda o, j, v: P[o, j].__iadd__(v)  # noqa: E731
    else:
        I = np.asarray(I, np.float64)
        out = np.zeros((3, len(s)))
        add = lambda o, j, v: out[o].__iadd__(I[j] * v)  # noqa: E731
    for j in range(E + 1):
        kj, dj = k[j], k[j + 1] - k[j]
        w = dj * np.exp(-kj * s)
        m0, m1, m2, m3 = _moments(dj * s)
        if j >= 1:
            c0, c1, c2 = m0 - m1, m1 - m2, m2 - m3
            add(0, j - 1, w * c0)
            add(1, j - 1, -w * (kj * c0 + dj * c1))
            add(2, j - 1, w * (kj * kj * c0 + 2 * kj * dj * c1 + dj * dj * c2))
        if j < E:
            add(0, j, w * m1)
            add(1, j, -w * (kj * m1 + dj * m2))
            add(2, j, w * (kj * kj * m1 + 2 * kj * dj * m2 + dj * dj * m3))
    if I is None:
        return P[0].T, P[1].T, P[2].T
    return out[0], out[1], out[2]


def b0_iout(k, s, I=None, nout=3):
    """bhc/Spline.m:205-262 (piecewise-constant basis)."""
    EPS = 1e-6
    k = np.asarray(k, np.float64).ravel()
    s = np.asarray(s, np.float64).ravel()
    E = len(k) - 1
    shape = (len(s),) if I is not None else (len(s), E)
    BL, sBL, ssBL = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        expksR = np.exp(-k[0] * s)
        s2, s3, abss = s ** 2, s ** 3, np.abs(s)
        for i in range(E):
            a, b = k[i], k[i + 1]
            expksL = expksR
            expksR = np.exp(-b * s)
            temp = (expksL - expksR) / s
            t1 = np.array([b - a, a ** 2 / 2 - b ** 2 / 2, b ** 3 / 6 - a ** 3 / 6, a ** 4 / 24 - b ** 4 / 24])
            thr = abs(EPS * t1[0] / t1[-1]) ** (1 / 3)
            idx = abss < thr
            temp[idx] = t1[0] + np.stack([s[idx], s2[idx], s3[idx]], 1) @ t1[1:]
            if I is not None:
                BL += temp * I[i]
            else:
                BL[:, i] = temp
            if nout > 1:
                temp = -((s * a + 1) * expksL - (s * b + 1) * expksR) / s2
                t1 = np.array([a ** 2 / 2 - b ** 2 / 2, b ** 3 / 3 - a ** 3 / 3, a ** 4 / 8 - b ** 4 / 8, b ** 5 / 30 - a ** 5 / 30])
                thr = abs(EPS * t1[0] / t1[-1]) ** (1 / 3)
                idx = abss < thr
                temp[idx] = t1[0] + np.stack([s[idx], s2[idx], s3[idx]], 1) @ t1[1:]
                if I is not None:
                    sBL += temp * I[i]
                else:
                    sBL[:, i] = temp
            if nout > 2:
                temp = (((a * s + 2) * a * s + 2) * expksL - ((b * s + 2) * b * s + 2) * expksR) / s3
                t1 = np.array([(b ** 3 - a ** 3) / 3, (a ** 4 - b ** 4) / 4, (b ** 5 - a ** 5) / 10, (a ** 6 - b ** 6) / 36])
                thr = abs(EPS * t1[0] / t1[-1]) ** (1 / 3)
                idx = abss < thr
                temp[idx] = t1[0] + t1[1] * s[idx] + t1[2] * s2[idx]
                if I is not None:
                    ssBL += temp * I[i]
                else:
                    ssBL[:, i] = temp
    return BL, sBL, ssBL


def dis_iout(k, s, I=None, nout=3):
    """bhc/Spline.m:264-302 (discrete kappa)."""
    k = np.asarray(k, np.float64).ravel()
    s = np.asarray(s, np.float64).ravel()
    dk = (k[2:] - k[:-2]) / 2
    kk = k[1:-1]
    shape = (len(s),) if I is not None else (len(s), len(kk))
    BL, sBL, ssBL = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    for i in range(len(kk)):
        temp = np.exp(-kk[i] * s) * dk[i]
        if I is not None:
            BL += temp * I[i]
        else:
            BL[:, i] = temp
        temp = temp * (-kk[i])
        if I is not None:
            sBL += temp * I[i]
        else:
            sBL[:, i] = temp
        temp = temp * (-kk[i])
        if I is not None:
            ssBL += temp * I[i]
        else:
            ssBL[:, i] = temp
    return BL, sBL, ssBL


def b0_iout_stable(k, s, I=None, nout=3):
    """The B0 (box) spectrum transform by segment moments, smooth in s: box j = [a, b],
    Delta = b - a, w = Delta exp(-a s), z = Delta s:
      phi = w M0,  phi' = -w (a M0 + Delta M1),  phi'' = w (a^2 M0 + 2 a Delta M1 + Delta^2 M2)
    (the kernels' formula, in FP64)."""
    k = np.asarray(k, np.float64).ravel()
    s = np.asarray(s, np.float64).ravel()
    E = len(k) - 1
    P = np.zeros((3, len(s), E))
    for j in range(E):
        a, d = k[j], k[j + 1] - k[j]
        w = d * np.exp(-a * s)
        M0, M1, M2, _ = _moments(d * s)
        P[0, :, j] = w * M0
        P[1, :, j] = -w * (a * M0 + d * M1)
        P[2, :, j] = w * (a * a * M0 + 2 * a * d * M1 + d * d * M2)
    if I is None:
        return P[0], P[1], P[2]
    I = np.asarray(I, np.float64)
    return P[0] @ I, P[1] @ I, P[2] @ I


def b0_iout_quad(k, s, j):
    """phi_j(s), phi_j'(s), phi_j''(s) of box j by adaptive quadrature (independent pin)."""
    from scipy.integrate import quad
    a, b = k[j], k[j + 1]
    f0 = quad(lambda x: np.exp(-x * s), a, b, epsabs=0, epsrel=1e-13, limit=200)[0]
    f1 = -quad(lambda x: x * np.exp(-x * s), a, b, epsabs=0, epsrel=1e-13, limit=200)[0]
    f2 = quad(lambda x: x * x * np.exp(-x * s), a, b, epsabs=0, epsrel=1e-13, limit=200)[0]
    return f0, f1, f2


def _nll(BL, Imea):
    """sum(BL-Imea) - Imea'*log(BL./Imea) with 0*log(.) = 0 for Imea = 0."""
    BL = np.asarray(BL, np.float64)
    Imea = np.asarray(Imea, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        err = np.where(Imea > 0, np.log(BL / np.where(Imea > 0, Imea, 1.0)), 0.0)
    return float(np.sum(BL - Imea) - Imea @ err)


def poiss_l_alpha(Imea, Ie, alpha, Phi, Phit, polyIout=b1_iout, knots=None):
    """bhc/poissLAlpha.m: f, g, Hessian ray weight."""
    R = Phi(alpha)
    BL, sBL, ssBL = polyIout(knots, R, Ie)
    f = _nll(BL, Imea)
    g = Phit((1 - Imea / BL) * sBL)
    hw = ((sBL / BL) ** 2) * Imea + ssBL * (1 - Imea / BL)
    return f, g, hw


def poiss_l_i(Imea, A, Ie):
    """bhc/poissLI.m: f, g = A'(1-Imea./Ir), H = A' diag(Imea./Ir.^2) A."""
    Ir = A @ np.asarray(Ie, np.float64)
    ratio = Imea / Ir
    g = A.T @ (1 - ratio)
    f = _nll(Ir, Imea)
    w = ratio / Ir
    H = A.T @ (A * w[:, None])
    return f, g, H


def gauss_l_alpha(Imea, Ie, alpha, Phi, Phit, polyIout=b1_iout, knots=None):
    """bhc/gaussLAlpha.m: Err = log(BL/Imea), f = Err'Err/2, g = Phit(Err.*sBL./BL),
    Hessian ray weight (1-Err).*(sBL./BL).^2 + Err.*ssBL./BL."""
    R = Phi(alpha)
    BL, sBL, ssBL = polyIout(knots, R, Ie)
    Err = np.log(BL / Imea)
    f = Err @ Err / 2
    g = Phit(Err * sBL / BL)
    hw = (1 - Err) * (sBL / BL) ** 2 + Err * ssBL / BL
    return f, g, hw


def gauss_l_i(Imea, A, Ie):
    """bhc/gaussLI.m: Ir = A Ie, Err = log(Ir./Imea), f = Err'Err/2, g = A'(Err./Ir),
    H = A' diag((1-Err)./Ir.^2) A."""
    Ir = A @ np.asarray(Ie, np.float64)
    Err = np.log(Ir / Imea)
    f = Err @ Err / 2
    g = A.T @ (Err / Ir)
    w = (1 - Err) / Ir ** 2
    H = A.T @ (A * w[:, None])
    return f, g, H


def ie_newton(Imea, A, Ie0, noise="poisson", maxit=50, tol=1e-14, trace=None):
    """Damped two-metric projected Newton minimisation of poissLI / gaussLI over Ie >= 0
    (Bertsekas' projected Newton with a Levenberg-Marquardt shift; the algorithm of
    ctprj_ie_solve, restated in FP64):
      w = |x - max(x - g, 0)|, eps = min(1e-2 max(x), w),
      c_j = max(|H_jj|, 1e-6 max_k |H_kk|)  (the diagonal scale; H may be indefinite:
      gaussLI's Hessian where Err > 1),
      active A = {x_j <= eps, g_j > 0}: d_A = g_A / c_A,
      free F: (H_FF + lam diag(c_F)) d_F = g_F (Cholesky; lam x10 while it fails),
      x(t) = max(x - t d, 0), t = 1, 1/2, ... until
      f(x(t)) <= f - 1e-4 (t g_F.d_F + g_A.(x_A - x_A(t)));
      lam /= 3 after a full step, x4 after a shortened one, x10 (same point) when no
      step decreases f.
    Stops when g_F.d_F + sum_A g_j min(x_j, d_j) <= tol max(|f|, 1), after maxit
    iterations, or when lam exceeds 1e12.  `trace`, if a list, receives f per accepted
    iterate."""
    obj = gauss_l_i if noise == "gaussian" else poiss_l_i
    x = np.maximum(np.asarray(Ie0, np.float64), 0.0)
    f, g, H = obj(Imea, A, x)
    lam = 1e-3
    if trace is not None:
        trace.append(f)
    for it in range(maxit):
        w = np.linalg.norm(x - np.maximum(x - g, 0.0))
        eps = min(1e-2 * x.max(), w)
        act = (x <= eps) & (g > 0)
        F = ~act
        hd = np.abs(np.diag(H))
        c = np.maximum(hd, 1e-6 * max(hd.max(), 1e-300))
        d = np.zeros_like(x)
        d[act] = g[act] / c[act]
        if F.any():
            M = H[np.ix_(F, F)]
            D = np.diag(c[F])
            while True:
                try:
                    L = np.linalg.cholesky(M + lam * D)
                    d[F] = np.linalg.solve(L.T, np.linalg.solve(L, g[F]))
                    break
                except np.linalg.LinAlgError:
                    lam *= 10
        gdF = g[F] @ d[F]
        dec = gdF + g[act] @ np.minimum(x[act], d[act])
        if not dec > tol * max(abs(f), 1.0):
            return x, f, it
        t, ok = 1.0, False
        for _ in range(30):
            xn = np.maximum(x - t * d, 0.0)
            fn = obj(Imea, A, xn)[0]
            if fn <= f - 1e-4 * (t * gdF + g[act] @ (x[act] - xn[act])):
                ok = True
                break
            t *= 0.5
        if not ok:
            lam *= 10
            if lam > 1e12:
                return x, f, it
            continue
        lam = max(lam / 3, 1e-12) if t == 1.0 else lam * 4
        x = xn
        f, g, H = obj(Imea, A, x)
        if trace is not None:
            trace.append(f)
    return x, f, maxit


def ie_minimize(Imea, A, Ie0, noise="poisson"):
    """Independent minimiser of poissLI / gaussLI over Ie >= 0: scipy's bounded
    L-BFGS-B with tight tolerances."""
    from scipy.optimize import minimize
    obj = gauss_l_i if noise == "gaussian" else poiss_l_i
    r = minimize(lambda v: obj(Imea, A, v)[:2], np.maximum(np.asarray(Ie0, np.float64), 0),
                 jac=True, method="L-BFGS-B", bounds=[(0, None)] * A.shape[1],
                 options={"ftol": 1e-15, "gtol": 1e-12, "maxiter": 20000, "maxcor": 30})
    return r.x, r.fun


def kappa_logspace(E, lo=0.03, hi=30.0):
    return np.logspace(np.log10(lo), np.log10(hi), E)


def spline_knots(kappa):
    k = np.asarray(kappa, np.float64)
    q = k[1] / k[0]
    return np.concatenate([[k[0] / q], k, [k[-1] * q]])


def energy_bins(epsilon):
    """Widths of the energy bins [eps_1, midpoints, eps_L] (bhc/genBeamHarden.m:53-54)."""
    eps = np.asarray(epsilon, np.float64).ravel()
    edges = np.concatenate([[eps[0]], (eps[:-1] + eps[1:]) / 2, [eps[-1]]])
    return edges[1:] - edges[:-1]


def gen_beam_harden(y, density, kappa, epsilon, iota, scale=2.0 ** -12 * 1.1):
    """bhc/genBeamHarden.m:53-79: y = K FP32 projections A x_k; returns (I / max I, Ts,
    max I) in FP64, evaluated in the reference's order (no fused multiply-adds)."""
    deps = energy_bins(epsilon)
    K = len(y)
    yk = [np.asarray(y[k], np.float32).astype(np.float64).ravel() * float(density[k])
          for k in range(K)]
    kap = np.asarray(kappa, np.float64).reshape(K, -1)
    io = np.asarray(iota, np.float64).ravel()
    Ts = 1.0
    while True:
        Imea = np.zeros_like(yk[0])
        for i in range(len(deps)):
            pa = np.zeros_like(yk[0])
            for k in range(K):
                pa = pa + yk[k] * kap[k, i]
            pa = pa * Ts
            Imea = Imea + np.exp(-pa) * io[i] * deps[i]
        temp = Imea.max()
        Imea = Imea / temp
        if Imea.min() < scale:
            Ts = Ts * 0.8
        else:
            return Imea, Ts, temp

