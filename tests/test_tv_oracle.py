"""The TV prox restatement (utils/denoise_bound_mod.m via utils/TV.m, utils/tlv.m,
utils/Lforward.m, utils/Ltrans.m) against identities and closed forms."""
import numpy as np
import pytest

from oracle import tv_oracle as T


def test_lforward_is_the_adjoint_of_ltrans():
    rng = np.random.default_rng(0)
    for m, n in [(2, 2), (7, 9), (16, 5)]:
        X = rng.standard_normal((m, n))
        P1, P2 = rng.standard_normal((m - 1, n)), rng.standard_normal((m, n - 1))
        q1, q2 = T.ltrans(X)
        lhs = (T.lforward(P1, P2) * X).sum()
        assert lhs == pytest.approx((P1 * q1).sum() + (P2 * q2).sum(), rel=1e-12)


def test_tlv_by_hand():
    X = np.array([[1.0, 4.0], [5.0, 2.0]])
    # vertical differences -4, 2; horizontal -3, 3
    assert T.tlv(X, "l1") == pytest.approx(4 + 2 + 3 + 3)
    assert T.tlv(X, "iso") == pytest.approx(np.hypot(4, 3) + 2 + 3)


@pytest.mark.parametrize("tv", ["iso", "l1"])
@pytest.mark.parametrize("lam", [0.3, 1.2])
def test_two_level_rows_closed_form(tv, lam):
    # equal rows [a, b]: each row is the 1-D problem with the step, prox = [a-lam, b+lam]
    a, b = 5.0, 2.0
    X = np.tile([a, b, b], (4, 1)) if tv == "l1" else np.tile([a, b], (3, 1))
    D, it = T.denoise(X, lam, tv, epsilon=1e-13, max_iter=20000)
    exp = np.tile([a - lam, b + lam / (2 if X.shape[1] == 3 else 1),
                   b + lam / 2][:X.shape[1]], (X.shape[0], 1))
    np.testing.assert_allclose(D, exp, atol=1e-8)


def test_constant_image_is_a_fixed_point():
    D, it = T.denoise(np.full((6, 7), 3.0), 0.7, "iso", epsilon=1e-10)
    np.testing.assert_allclose(D, 3.0, atol=1e-12)


@pytest.mark.parametrize("tv", ["iso", "l1"])
def test_prox_optimality(tv):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((12, 10)) + 0.3
    lam = 0.4
    D, _ = T.denoise(X, lam, tv, epsilon=1e-12, max_iter=50000)
    assert D.min() >= 0
    f = lambda Z: 0.5 * ((Z - X) ** 2).sum() + lam * T.tlv(Z, tv)  # noqa: E731
    f0 = f(D)
    for _ in range(30):
        Z = np.maximum(D + 1e-4 * rng.standard_normal(D.shape), 0)
        assert f(Z) >= f0 - 1e-9
    assert f0 <= f(np.maximum(X, 0))
