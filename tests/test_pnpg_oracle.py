"""Pins the FP64 PNPG restatement (oracle/pnpg_oracle.py, npg/PNPG.m with the
non-negativity prox): on 0.5||Ax-b||^2 over x >= 0 it must reach the non-negative
least-squares solution (scipy.optimize.nnls), decrease the objective along accepted
iterates, and exercise rejects (a step that is too long at the start)."""
import numpy as np
from scipy.optimize import nnls

from oracle.pnpg_oracle import PnpgOracle


def _problem(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((80, 30))
    x_true = np.maximum(rng.standard_normal(30), 0)
    b = A @ x_true + 0.3 * rng.standard_normal(80)
    return A, b


def test_converges_to_nnls():
    A, b = _problem(1)

    def func(x, grad=True):
        r = A @ x - b
        return 0.5 * r @ r, (A.T @ r if grad else None)

    L = np.linalg.norm(A, 2) ** 2
    o = PnpgOracle(func, np.zeros(30), t0=L / 50)  # start with a too long step
    for _ in range(600):
        o.trial()
    x_star, _ = nnls(A, b)
    assert np.linalg.norm(o.alpha - x_star) <= 1e-6 * max(1.0, np.linalg.norm(x_star))
    kinds = [k for k, _ in o.log]
    assert "reject" in kinds and "accept" in kinds
    acc = [f for k, f in o.log if k == "accept"]
    assert acc[-1] <= acc[0]


def test_restart_keeps_monotone_objective():
    A, b = _problem(2)

    def func(x, grad=True):
        r = A @ x - b
        return 0.5 * r @ r, (A.T @ r if grad else None)

    o = PnpgOracle(func, np.ones(30), t0=np.linalg.norm(A, 2) ** 2)
    for _ in range(300):
        o.trial()
    acc = [f for k, f in o.log if k == "accept"]
    assert all(b2 <= b1 * (1 + 1e-12) for b1, b2 in zip(acc, acc[1:]))
