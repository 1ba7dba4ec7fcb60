"""FP64 NumPy restatement of the total-variation proximal step -- TEST INFRASTRUCTURE ONLY.

Follows, line by line:
  * ``ltrans``   -- utils/Ltrans.m   (P1 = -diff(X, 1, 1), P2 = -diff(X, 1, 2))
  * ``lforward`` -- utils/Lforward.m (its adjoint)
  * ``tlv``      -- utils/tlv.m      (isotropic or l1 total variation)
  * ``denoise``  -- utils/denoise_bound_mod.m with the bounds [0, inf) that TV.denoise
                    passes (utils/TV.m:56-80): FGP (fast gradient projection on the
                    dual), stopping after MAXITER iterations or five in a row with
                    ||D - Dold|| / ||D|| < epsilon.  TV.denoise never forwards its `init`
                    (utils/TV.m:62 sets pars.init only when init is absent), so the dual
                    variables start at zero.

Pinned by: the adjoint identity <Lforward(P), X> = <P, Ltrans(X)>, closed forms of the
prox on images made of equal rows (a 1-D two-level step), and the prox optimality (no
feasible perturbation lowers the primal objective).
"""
from __future__ import annotations

import numpy as np


def ltrans(X):
    return -np.diff(X, 1, 0), -np.diff(X, 1, 1)


def lforward(P1, P2):
    m, n = P1.shape[0] + 1, P2.shape[1] + 1
    X = np.zeros((m, n))
    X[0, :] = P1[0, :]
    X[-1, :] = -P1[-1, :]
    X[1:-1, :] = np.diff(P1, 1, 0)
    Y = np.zeros((m, n))
    Y[:, 0] = P2[:, 0]
    Y[:, -1] = -P2[:, -1]
    Y[:, 1:-1] = np.diff(P2, 1, 1)
    return X + Y


def tlv(X, tv="iso"):
    m, n = X.shape
    P1, P2 = ltrans(X)
    if tv == "iso":
        D = np.zeros((m, n))
        D[:m - 1, :] = P1 ** 2
        D[:, :n - 1] = D[:, :n - 1] + P2 ** 2
        return float(np.sqrt(D).sum())
    return float(np.abs(P1).sum() + np.abs(P2).sum())


def _project_dual(P1, P2, tv):
    if tv == "iso":
        m, n = P1.shape[0] + 1, P2.shape[1] + 1
        A = np.vstack([P1, np.zeros((1, n))]) ** 2 + np.hstack([P2, np.zeros((m, 1))]) ** 2
        A = np.sqrt(np.maximum(A, 1))
        return P1 / A[:m - 1, :], P2 / A[:, :n - 1]
    return np.clip(P1, -1, 1), np.clip(P2, -1, 1)


def denoise(Xobs, lam, tv="iso", epsilon=1e-4, max_iter=100):
    """argmin 0.5||D - Xobs||^2 + lam TV(D) over D >= 0; returns (D, iterations)."""
    Xobs = np.asarray(Xobs, np.float64)
    m, n = Xobs.shape
    D = np.zeros((m, n))
    P1, P2 = ltrans(D)
    P1, P2 = _project_dual(P1, P2, tv)
    R1, R2 = P1, P2
    tk = tkp1 = 1.0
    count = 0
    i = 0
    fval = np.inf
    while i < max_iter and count < 5:
        fold = fval
        i += 1
        Dold = D
        Pold1, Pold2 = P1, P2
        tk = tkp1
        C = Xobs - lam * lforward(R1, R2)
        D = np.maximum(C, 0)
        Q1, Q2 = ltrans(D)
        P1 = R1 + Q1 / (8 * lam)
        P2 = R2 + Q2 / (8 * lam)
        P1, P2 = _project_dual(P1, P2, tv)
        tkp1 = (1 + np.sqrt(1 + 4 * tk ** 2)) / 2
        R1 = P1 + (tk - 1) * (P1 - Pold1) / tkp1
        R2 = P2 + (tk - 1) * (P2 - Pold2) / tkp1
        re = np.sqrt(((D - Dold) ** 2).sum()) / np.sqrt((D ** 2).sum())
        count = count + 1 if re < epsilon else 0
        fval = -((C - D) ** 2).sum() + (C ** 2).sum()
        if fval > fold:
            tkp1 = 1.0
    return D, i
