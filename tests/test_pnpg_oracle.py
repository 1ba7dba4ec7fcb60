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



DECIDE_CASES = [
    # (cost, f_old, f_new, ls, moved, admmTol): expected code
    ((10.0, 9.0, 9.5, 1, 1, 1e-2), 2),    # lower objective: take newX
    ((10.0, 9.0, 10.5, 1, 1, 1e-2), 0),   # higher, momentum moved xbar: restart
    ((10.0, 11.0, 10.5, 1, 0, 1e-2), 2),  # higher, no momentum, objBar >= newObj: take
    ((10.0, 9.0, 10.5, 1, 0, 1e-2), 1),   # higher, no momentum, objBar < newObj: retry
    ((10.0, 9.0, 10.5, 1, 0, 1e-6), 3),   # ... with admmTol used up: keep alpha
    ((10.0, 12.0, 10.5, 2, 1, 1e-3), 1),  # given up, higher: reset, retry
    ((10.0, 12.0, 10.5, 2, 1, 1e-7), 3),  # given up, admmTol used up: keep alpha
    ((10.0, 12.0, 9.0, 2, 1, 1e-2), 2),   # given up but lower: take newX
]


def test_decision_rule_restatement():
    """PNPG.m:99-160 (u = 0): the restart needs momentum (xbar != alpha), a higher
    objective without it retries through the admmTol schedule and then keeps alpha, and a
    given-up line search resets theta."""
    from oracle.pnpg_oracle import PnpgOracle
    for (cost, fo, fn, ls, moved, tol), want in DECIDE_CASES:
        o = PnpgOracle(lambda x, grad=True: (0.0, None), np.zeros(3), 1.0)
        o.cost, o.admmTol, o.theta = cost, tol, 0.7
        code = o._decide(ls, fo, fn, bool(moved))
        assert code == want, ((cost, fo, fn, ls, moved, tol), code)
        if ls == 2 and code in (1, 3):
            assert o.theta == 1.0
        if code == 1:
            assert o.admmTol == pytest.approx(tol / 10)


def test_device_decision_rule_matches_restatement():
    """The library's decision function (the one k_pnpg_decide runs, called on the host)
    gives the restatement's outcome, theta and admmTol on every case."""
    import ctypes
    import os
    from imgrecsrc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libctprj.so not built")
    from oracle.pnpg_oracle import PnpgOracle
    L = _lib.lib()
    for (cost, fo, fn, ls, moved, tol), want in DECIDE_CASES:
        inp = (ctypes.c_double * 5)(cost, fo, fn, 0.7, tol)
        out = (ctypes.c_double * 3)()
        code = ctypes.c_int(-1)
        _lib.check(L.ctprj_debug_pnpg_decide(inp, ls, moved, out, ctypes.byref(code)))
        o = PnpgOracle(lambda x, grad=True: (0.0, None), np.zeros(3), 1.0)
        o.cost, o.admmTol, o.theta = cost, tol, 0.7
        ref = o._decide(ls, fo, fn, bool(moved))
        assert code.value == ref == want
        assert out[2] == pytest.approx(o.admmTol)
        if ref in (0, 1):
            assert out[1] == o.theta
