"""The wavelet restatement (rwt/mdwt_r.c, rwt/midwt_r.c, npg/admm.m) and the host-side
filter design (rwt/daubcqf.m): against the reference's own transforms compiled into
oracle/_ref, published Daubechies coefficients, the orthonormality conditions, and the
closed form of the prox where the non-negativity constraint is inactive."""
import numpy as np
import pytest

from oracle import wavelet_oracle as W
from imgrecsrc_amd.wavelet import daubcqf

# Daubechies (1988) minimum-phase scaling filters, normalised to sum sqrt(2)
D4 = np.array([1 + np.sqrt(3), 3 + np.sqrt(3), 3 - np.sqrt(3), 1 - np.sqrt(3)]) / (4 * np.sqrt(2))
D6 = np.array([0.33267055295008263, 0.80689150931109255, 0.45987750211849154,
               -0.13501102001025458, -0.08544127388202666, 0.03522629188570953])


def test_daubcqf_known_filters():
    np.testing.assert_allclose(daubcqf(2)[0], [2 ** -0.5, 2 ** -0.5], rtol=1e-15)
    np.testing.assert_allclose(daubcqf(4)[0], D4, rtol=1e-13)
    np.testing.assert_allclose(daubcqf(6)[0], D6, rtol=1e-12)
    np.testing.assert_allclose(daubcqf(4, "max")[0], D4[::-1], rtol=1e-13)
    # the wavelet filter: reversed scaling filter with alternating signs (rwt/daubcqf.m)
    h0, h1 = daubcqf(6)
    np.testing.assert_allclose(h1, h0[::-1] * np.array([-1, 1, -1, 1, -1, 1]), rtol=1e-15)


@pytest.mark.parametrize("N", [2, 4, 6, 8, 10, 12, 16, 20])
def test_daubcqf_orthonormal_with_vanishing_moments(N):
    h = daubcqf(N)[0]
    assert h.sum() == pytest.approx(np.sqrt(2), rel=1e-13)
    for m in range(N // 2):  # sum_k h_k h_{k+2m} = delta_m
        assert float(h[:N - 2 * m] @ h[2 * m:]) == pytest.approx(1.0 if m == 0 else 0.0,
                                                                 abs=1e-11)
    k = np.arange(N)
    g = (-1.0) ** k * h[::-1]
    for p in range(N // 2):  # N/2 vanishing moments of the wavelet
        assert abs(float(((k / N) ** p) @ g)) < 1e-9
    # minimum phase: the zeros of H(z) other than z = -1 lie inside the unit circle (the
    # N/2-fold zero at -1 spreads numerically beyond N = 12)
    if N <= 12:
        r = np.roots(h)
        assert np.all(np.abs(r[np.abs(r + 1) > 0.05]) < 1.0)


CASES = [(2, 3, (16, 16)), (4, 3, (32, 16)), (6, 4, (64, 64)), (8, 2, (8, 12)),
         (6, 5, (32, 32)), (20, 2, (16, 16)), (4, 1, (2, 6))]


@pytest.mark.skipif(not W.ref_available(), reason="oracle/_ref/librwt_ref.so not built")
@pytest.mark.parametrize("N,L,shape", CASES)
def test_transforms_match_the_reference_bit_for_bit(N, L, shape):
    h = daubcqf(N)[0]
    x = np.random.default_rng(N * 10 + L).standard_normal(shape)
    y = W.mdwt(x, h, L)
    np.testing.assert_array_equal(y, W.ref_mdwt(x, h, L))
    np.testing.assert_array_equal(W.midwt(y, h, L), W.ref_midwt(y, h, L))


@pytest.mark.parametrize("N,L,shape", CASES)
def test_transforms_are_orthonormal(N, L, shape):
    h = daubcqf(N)[0]
    x = np.random.default_rng(7).standard_normal(shape)
    y = W.mdwt(x, h, L)
    assert (y ** 2).sum() == pytest.approx((x ** 2).sum(), rel=1e-11)
    np.testing.assert_allclose(W.midwt(y, h, L), x, atol=1e-10)


def test_haar_level_by_hand():
    h = daubcqf(2)[0]
    x = np.array([[1.0, 3.0], [5.0, 11.0]])
    y = W.mdwt(x, h, 1)
    # rows: lo = (a + b)/sqrt2, hi = (a - b)/sqrt2 ... then columns
    r = np.array([[4.0, -2.0], [16.0, -6.0]]) / np.sqrt(2)
    c = np.array([[r[0, 0] + r[1, 0], r[0, 1] + r[1, 1]],
                  [r[0, 0] - r[1, 0], r[0, 1] - r[1, 1]]]) / np.sqrt(2)
    np.testing.assert_allclose(y, c, rtol=1e-15)


def test_admm_matches_the_closed_form_when_nonnegativity_is_inactive():
    h = daubcqf(4)[0]
    L = 3
    rng = np.random.default_rng(3)
    a = 10.0 + rng.standard_normal((32, 32))  # far from the bound
    u = 0.5
    x, itr = W.admm(a, u, h, L, relative_tol=1e-10, max_itr=5000)
    closed = W.midwt(W.soft_thresh(W.mdwt(a, h, L), u), h, L)
    assert closed.min() > 0
    np.testing.assert_allclose(x, closed, atol=1e-7 * np.abs(a).max())
    assert itr > 1


def test_admm_satisfies_the_prox_optimality_conditions():
    # x = prox(a): with v = W x, 0 in x - a + W^T d + n, d in u d|v|_1, n in N_{x>=0}
    h = daubcqf(6)[0]
    L = 2
    rng = np.random.default_rng(4)
    a = rng.standard_normal((16, 16))
    u = 0.3
    x, _ = W.admm(a, u, h, L, relative_tol=1e-12, max_itr=20000)
    assert x.min() >= 0 and np.any(x == 0)
    # on the free set (x > 0) the gradient a - x is in u W^T d|Wx|_1
    # => for the primal objective, small feasible perturbations do not decrease it
    f = lambda z: 0.5 * ((z - a) ** 2).sum() + u * np.abs(W.mdwt(z, h, L)).sum()  # noqa: E731
    f0 = f(x)
    for k in range(20):
        d = rng.standard_normal(a.shape) * 1e-4
        z = np.maximum(x + d, 0)
        assert f(z) >= f0 - 1e-9
    # and it beats the nonneg projection of the unconstrained solution
    assert f0 <= f(np.maximum(W.midwt(W.soft_thresh(W.mdwt(a, h, L), u), h, L), 0)) + 1e-12


def test_admm_of_zero_is_zero():
    x, itr = W.admm(np.zeros((8, 8)), 1.0, daubcqf(2)[0], 2)
    assert itr == 0 and not x.any()


def _disc(n, r):
    c = np.arange(n) - (n - 1) / 2
    X, Y = np.meshgrid(c, c)
    return X ** 2 + Y ** 2 <= r * r


@pytest.mark.parametrize("N,L,n", [(2, 2, 16), (4, 3, 32), (6, 2, 16)])
def test_wvlt_mask_host_helper_matches_the_definition(N, L, n):
    from imgrecsrc_amd.wavelet import wvlt_mask
    h = daubcqf(N)[0]
    m = _disc(n, n / 3)
    np.testing.assert_array_equal(wvlt_mask(m, h, L), W.wvlt_mask(m, h, L))


def test_masked_admm_is_the_prox_on_the_support():
    # the masked ADMM works on the disc's vector and satisfies the prox optimality there
    h = daubcqf(4)[0]
    L, n = 2, 16
    m = _disc(n, 6)
    mk = W.wvlt_mask(m, h, L)
    a = (10.0 + np.random.default_rng(2).standard_normal((n, n)))[m]
    ops = W.masked_ops(h, L, m, mk)
    x, itr = W.admm(a, 0.3, h, L, 1e-11, 20000, ops=ops)
    assert x.shape == a.shape and itr > 1
    f = lambda z: 0.5 * ((z - a) ** 2).sum() + 0.3 * np.abs(ops[1](z)).sum()  # noqa: E731
    rng = np.random.default_rng(3)
    for _ in range(20):
        z = np.maximum(x + 1e-4 * rng.standard_normal(x.shape), 0)
        assert f(z) >= f(x) - 1e-9
