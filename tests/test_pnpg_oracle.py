"""Pins the FP64 PNPG / NPG restatements (oracle/pnpg_oracle.py, npg/PNPG.m and
npg/NPG.m with the non-negativity prox): on 0.5||Ax-b||^2 over x >= 0 they must reach the
non-negative least-squares solution (scipy.optimize.nnls), decrease the objective along
accepted iterates, and exercise rejects (a step that is too long at the start); NPG's
first iterations grow a too short step.  With the wavelet prior the restatement must
reach the minimiser of 0.5||Ax-b||^2 + u||W x||_1 over x >= 0 found by an independent
method (a long run of plain proximal gradient steps with a tightly solved prox)."""
import numpy as np
import pytest
from scipy.optimize import nnls

from oracle.pnpg_oracle import PnpgOracle


def _problem(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((80, 30))
    x_true = np.maximum(rng.standard_normal(30), 0)
    b = A @ x_true + 0.3 * rng.standard_normal(80)
    return A, b


@pytest.mark.parametrize("npg", [False, True])
def test_converges_to_nnls(npg):
    A, b = _problem(1)

    def func(x, grad=True):
        r = A @ x - b
        return 0.5 * r @ r, (A.T @ r if grad else None)

    L = np.linalg.norm(A, 2) ** 2
    o = PnpgOracle(func, np.zeros(30), t0=L / 50, npg=npg)  # start with a too long step
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


def test_npg_grows_a_short_step():
    A, b = _problem(3)

    def func(x, grad=True):
        r = A @ x - b
        return 0.5 * r @ r, (A.T @ r if grad else None)

    L = np.linalg.norm(A, 2) ** 2
    o = PnpgOracle(func, np.zeros(30), t0=8 * L, npg=True)  # step 1/t far too short
    for _ in range(12):
        o.trial()
    kinds = [k for k, _ in o.log]
    assert kinds.count("grow") >= 3 and o.t < 8 * L / 4


def test_wavelet_prior_reaches_the_regularised_minimum():
    from oracle.pnpg_oracle import PnpgWaveletOracle
    from oracle import wavelet_oracle as W
    from imgrecsrc_amd.wavelet import daubcqf
    rng = np.random.default_rng(6)
    shape = (8, 8)
    A = rng.standard_normal((96, 64))
    b = A @ np.maximum(rng.standard_normal(64), 0) + 0.2 * rng.standard_normal(96)
    h, L = daubcqf(4)[0], 2
    u = 2.0

    def func(x, grad=True):
        r = A @ x - b
        return 0.5 * r @ r, (A.T @ r if grad else None)

    Lip = np.linalg.norm(A, 2) ** 2
    o = PnpgWaveletOracle(func, np.zeros(64), Lip, u, h, L, shape)
    for _ in range(110):
        o.trial()
    # independent: ISTA with step 1/Lip and the (tightly solved) prox
    x = np.zeros(64)
    for _ in range(600):
        g = A.T @ (A @ x - b)
        x = W.admm((x - g / Lip).reshape(shape), u / Lip, h, L, 1e-9, 5000)[0].ravel()
    F = lambda z: func(z, False)[0] + u * W.l1_penalty(z.reshape(shape), h, L)  # noqa: E731
    assert F(o.alpha) <= F(x) * (1 + 1e-9)
    assert np.linalg.norm(o.alpha - x) <= 1e-6 * np.linalg.norm(x)

