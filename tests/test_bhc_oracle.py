"""Pins the FP64 beam-hardening restatement (oracle/bhc_oracle.py) -- the reference's
BHC code is MATLAB and cannot run here -- and the moment form of the B1 spectrum
transform used by the HIP kernels:

* phi_j, phi_j', phi_j'' against adaptive quadrature of the hat-function Laplace
  transform, and phi_j(0) = (k_{j+2} - k_j)/2 in closed form;
* the Taylor branch of bhc/Spline.m:156-165 continuous at its threshold;
* poissLAlpha / poissLI gradients and Hessians against central finite differences
  (h = 1e-3 as in the workload definition, and smaller), and the identity
  g . Ie = sum(Ir - Imea) of poissLI; gradient 0 at Imea = Ir;
* the box (b0) basis against quadrature of the box Laplace transform, its reference
  form against the moment form; the discrete (dis) basis' derivatives against finite
  differences and its value at s = 0;
* gaussLAlpha / gaussLI gradients and Hessians against finite differences, gradient 0
  at Imea = Ir."""
import numpy as np
import pytest

from oracle import bhc_oracle as B

KN = B.spline_knots(B.kappa_logspace(20, 0.03, 30.0))
S = np.concatenate([[0.0, 1e-9, 1e-6, 1e-4], np.logspace(-3, 1.5, 40)])


def test_basis_at_zero():
    BL, _, _ = B.b1_iout(KN, np.array([0.0]), None)
    assert np.allclose(BL[0], (KN[2:] - KN[:-2]) / 2, rtol=1e-14)


@pytest.mark.parametrize("j", [0, 5, 10, 19])
def test_b1_against_quadrature(j):
    BL, sBL, ssBL = B.b1_iout(KN, S, None)
    for i, s in enumerate(S):
        q0, q1, q2 = B.b1_iout_quad(KN, s, j)
        # the reference's own accuracy target: 1e-6 of the value at s = 0 (Spline.m:127)
        assert abs(BL[i, j] - q0) <= 2e-6 * BL[0, j] + 1e-12 * abs(q0)
        assert abs(sBL[i, j] - q1) <= 2e-6 * abs(sBL[0, j]) + 1e-9 * abs(q1)
        assert abs(ssBL[i, j] - q2) <= 2e-6 * abs(ssBL[0, j]) + 1e-9 * abs(q2)


@pytest.mark.parametrize("s", [0.0, 1e-7, 1e-3, 0.05, 0.4, 2.0, 9.0, 30.0, -0.01])
def test_moment_form_against_quadrature(s):
    """The smooth moment form (the kernels' formula, here in FP64) against quadrature."""
    got = B.b1_iout_stable(KN, np.array([s]), None)
    for j in range(0, 20, 3):
        q = B.b1_iout_quad(KN, s, j)
        for p in range(3):
            assert abs(got[p][0, j] - q[p]) <= 1e-10 * max(abs(q[p]), 1e-300) + 1e-300, (s, j, p)


def test_reference_form_matches_moment_form():
    """bhc/Spline.m's closed form + Taylor branch agrees with the smooth form to its own
    accuracy target (1e-6 of phi(0)) over the whole s range."""
    a = B.b1_iout(KN, S, None)
    b = B.b1_iout_stable(KN, S, None)
    for p in range(3):
        scale = np.abs(b[p][0])
        assert np.all(np.abs(a[p] - b[p]) <= 2e-6 * scale + 1e-9 * np.abs(b[p]))


def test_taylor_branch_continuous():
    """Values just below and above each basis' switch point agree to the reference's
    1e-6-of-phi(0) accuracy target."""
    for j in range(20):
        a, b, c = KN[j], KN[j + 1], KN[j + 2]
        t1 = np.array([-a / 2 + c / 2,
                       (a ** 3 - b ** 3) / (6 * (a - b)) - (b ** 3 - c ** 3) / (6 * (b - c)),
                       -(a ** 4 - b ** 4) / (24 * (a - b)) + (b ** 4 - c ** 4) / (24 * (b - c)),
                       (a ** 5 - b ** 5) / (120 * (a - b)) - (b ** 5 - c ** 5) / (120 * (b - c))])
        thr = abs(1e-6 * t1[0] / t1[-1]) ** (1 / 3)
        ss = np.array([thr * (1 - 1e-9), thr * (1 + 1e-9)])
        BL, _, _ = B.b1_iout(KN, ss, None)
        assert abs(BL[0, j] - BL[1, j]) <= 2e-6 * t1[0]


def _toy_problem(seed=0, rays=300, pix=40):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0, 1, (rays, pix)) * (rng.uniform(0, 1, (rays, pix)) < 0.2)
    alpha = rng.uniform(0.0, 0.2, pix)
    Ie = np.exp(-0.5 * ((np.arange(20) - 8) / 3.0) ** 2)
    Ie = Ie / ((KN[2:] - KN[:-2]) / 2 @ Ie)
    s = A @ alpha
    BL, _, _ = B.b1_iout_stable(KN, s, Ie)
    Imea = rng.poisson(1e5 * BL) / 1e5
    # evaluate away from the truth: at the optimum g ~ 0 and relative FD errors blow up
    alpha = alpha * rng.uniform(0.5, 1.5, pix)
    return A, alpha, Ie, Imea


def central(fun, h):
    """Central difference at step h and its Richardson extrapolation with h/2
    (truncation O(h^2) -> O(h^4))."""
    d1 = (fun(h) - fun(-h)) / (2 * h)
    d2 = (fun(h / 2) - fun(-h / 2)) / h
    return d1, (4 * d2 - d1) / 3


def test_poiss_l_alpha_finite_differences():
    A, alpha, Ie, Imea = _toy_problem()
    Phi, Phit = (lambda x: A @ x), (lambda y: A.T @ y)
    # the smooth form: bhc/Spline.m's branch switch jumps by up to 1e-6 phi(0), which
    # would dominate a finite difference
    pol = lambda k, s, I: B.b1_iout_stable(k, s, I)  # noqa: E731
    f, g, hw = B.poiss_l_alpha(Imea, Ie, alpha, Phi, Phit, pol, KN)
    rng = np.random.default_rng(1)
    for _ in range(5):
        v = rng.standard_normal(alpha.size)
        fun = lambda t, v=v: B.poiss_l_alpha(Imea, Ie, alpha + t * v, Phi, Phit, pol, KN)[0]  # noqa: E731
        d1, rich = central(fun, 1e-3)
        gv = g @ v
        assert abs(d1 - gv) <= 1e-3 * abs(gv)       # O(h^2) truncation at h = 1e-3
        assert abs(rich - gv) <= 1e-7 * abs(gv)
        d5, _ = central(fun, 1e-5)
        assert abs(d5 - gv) <= 1e-6 * abs(gv)
        # Hessian along v: v' Phi' diag(hw) Phi v vs the second difference
        h = 1e-3
        fdd = (fun(h) - 2 * f + fun(-h)) / h ** 2
        hv = (A @ v) @ (hw * (A @ v))
        assert abs(fdd - hv) <= 1e-3 * abs(hv)


def test_poiss_l_i_finite_differences_and_identities():
    A, alpha, Ie, Imea = _toy_problem(seed=2)
    Amat, _, _ = B.b1_iout_stable(KN, A @ alpha, None)
    f, g, H = B.poiss_l_i(Imea, Amat, Ie)
    h = 1e-3
    for j in range(20):
        e = np.zeros(20)
        e[j] = 1.0
        _, rich = central(lambda t: B.poiss_l_i(Imea, Amat, Ie + t * e)[0], h)
        assert abs(rich - g[j]) <= 1e-7 * np.abs(g).max()
        _, gd = central(lambda t: B.poiss_l_i(Imea, Amat, Ie + t * e)[1], h)
        assert np.allclose(gd, H[:, j], rtol=1e-6, atol=1e-9 * np.abs(H).max())
    Ir = Amat @ Ie
    assert g @ Ie == pytest.approx(np.sum(Ir - Imea), rel=1e-10, abs=1e-10)
    _, g0, _ = B.poiss_l_i(Ir, Amat, Ie)
    assert np.abs(g0).max() < 1e-10 * np.abs(Amat).sum()


def test_zero_counts_guard():
    Amat = np.abs(np.random.default_rng(3).standard_normal((50, 20)))
    Ie = np.ones(20)
    Imea = np.abs(np.random.default_rng(4).standard_normal(50))
    Imea[::7] = 0.0
    f, g, H = B.poiss_l_i(Imea, Amat, Ie)
    assert np.isfinite(f) and np.all(np.isfinite(g)) and np.all(np.isfinite(H))


KN0 = np.logspace(np.log10(0.03), np.log10(30.0), 21)   # b0: 21 edges -> 20 boxes


@pytest.mark.parametrize("s", [0.0, 1e-7, 1e-3, 0.05, 0.4, 2.0, 9.0, 30.0, -0.01])
def test_b0_against_quadrature(s):
    stable = B.b0_iout_stable(KN0, np.array([s]), None)
    ref = B.b0_iout(KN0, np.array([s]), None)
    ref0 = B.b0_iout(KN0, np.array([0.0]), None)
    for j in range(0, 20, 3):
        q = B.b0_iout_quad(KN0, s, j)
        for p in range(3):
            assert abs(stable[p][0, j] - q[p]) <= 1e-10 * abs(q[p]) + 1e-300, (s, j, p)
            # the reference's closed form + Taylor branch: its own 1e-6-of-phi(0) target
            assert abs(ref[p][0, j] - q[p]) <= 2e-6 * abs(ref0[p][0, j]) + 1e-9 * abs(q[p])
    assert np.allclose(B.b0_iout_stable(KN0, np.array([0.0]), None)[0][0], np.diff(KN0),
                       rtol=1e-14)


def test_dis_basis_derivatives_and_zero():
    s = np.logspace(-3, 1, 30)
    BL, sBL, ssBL = B.dis_iout(KN, s, None)
    h = 1e-6
    Bp = B.dis_iout(KN, s + h, None)
    Bm = B.dis_iout(KN, s - h, None)
    assert np.allclose((Bp[0] - Bm[0]) / (2 * h), sBL, rtol=1e-6, atol=1e-12)
    assert np.allclose((Bp[1] - Bm[1]) / (2 * h), ssBL, rtol=1e-6, atol=1e-12)
    BL0, _, _ = B.dis_iout(KN, np.array([0.0]), None)
    assert np.allclose(BL0[0], (KN[2:] - KN[:-2]) / 2, rtol=1e-14)


def test_gauss_l_alpha_finite_differences():
    A, alpha, Ie, Imea = _toy_problem(seed=5)
    Phi, Phit = (lambda x: A @ x), (lambda y: A.T @ y)
    pol = lambda k, s, I: B.b1_iout_stable(k, s, I)  # noqa: E731
    f, g, hw = B.gauss_l_alpha(Imea, Ie, alpha, Phi, Phit, pol, KN)
    rng = np.random.default_rng(6)
    for _ in range(5):
        v = rng.standard_normal(alpha.size)
        fun = lambda t, v=v: B.gauss_l_alpha(Imea, Ie, alpha + t * v, Phi, Phit, pol, KN)[0]  # noqa: E731
        _, rich = central(fun, 1e-3)
        gv = g @ v
        assert abs(rich - gv) <= 1e-7 * abs(gv)
        h = 1e-3
        fdd = (fun(h) - 2 * f + fun(-h)) / h ** 2
        hv = (A @ v) @ (hw * (A @ v))
        assert abs(fdd - hv) <= 1e-3 * abs(hv)
    # a planted sign error in the Hessian weight is caught by the same check
    bad = -hw
    v = rng.standard_normal(alpha.size)
    fun = lambda t: B.gauss_l_alpha(Imea, Ie, alpha + t * v, Phi, Phit, pol, KN)[0]  # noqa: E731
    fdd = (fun(1e-3) - 2 * f + fun(-1e-3)) / 1e-6
    assert abs(fdd - (A @ v) @ (bad * (A @ v))) > 0.5 * abs(fdd)


def test_gauss_l_i_finite_differences():
    A, alpha, Ie, Imea = _toy_problem(seed=7)
    Amat, _, _ = B.b1_iout_stable(KN, A @ alpha, None)
    f, g, H = B.gauss_l_i(Imea, Amat, Ie)
    for j in range(20):
        e = np.zeros(20)
        e[j] = 1.0
        _, rich = central(lambda t: B.gauss_l_i(Imea, Amat, Ie + t * e)[0], 1e-3)
        assert abs(rich - g[j]) <= 1e-7 * np.abs(g).max()
        _, gd = central(lambda t: B.gauss_l_i(Imea, Amat, Ie + t * e)[1], 1e-3)
        assert np.allclose(gd, H[:, j], rtol=1e-6, atol=1e-9 * np.abs(H).max())
    Ir = Amat @ Ie
    f0, g0, _ = B.gauss_l_i(Ir, Amat, Ie)
    assert f0 == 0.0 and np.abs(g0).max() == 0.0


def _ie_problem(seed, noise, E=20):
    """Poisson or log-normal data of a spectrum with empty bins, so the optimum has
    active bounds.  E = 20: the realistic, severely ill-conditioned basis matrix
    (a discretised Laplace transform); E = 6 on a coarse grid: well conditioned."""
    rng = np.random.default_rng(seed)
    kn = B.spline_knots(B.kappa_logspace(E, 0.03, 30.0))
    s = rng.uniform(0.0, 3.0, 400)
    Amat, _, _ = B.b1_iout_stable(kn, s, None)
    truth = np.exp(-0.5 * ((np.arange(E) - E / 2) / (E / 6)) ** 2)
    truth[:E // 5] = 0.0
    truth[-max(E // 6, 1):] = 0.0
    truth /= ((kn[2:] - kn[:-2]) / 2) @ truth
    Ir = Amat @ truth
    Imea = (rng.poisson(1e5 * Ir) / 1e5 if noise == "poisson"
            else Ir * np.exp(0.01 * rng.standard_normal(Ir.size)))
    return Amat, Imea, np.full(E, 1.0 / E) / (((kn[2:] - kn[:-2]) / 2).sum() / E)


@pytest.mark.parametrize("noise", ["poisson", "gaussian"])
def test_ie_newton_well_conditioned_reaches_the_minimiser(noise):
    """On a well-conditioned problem the damped projected Newton restatement (the
    library's I-epsilon step) and scipy's bounded L-BFGS-B reach the same minimum, with
    the bounds that the data leave empty active and the KKT conditions satisfied."""
    Amat, Imea, Ie0 = _ie_problem(11, noise, E=6)
    obj = B.gauss_l_i if noise == "gaussian" else B.poiss_l_i
    x, f, its = B.ie_newton(Imea, Amat, Ie0, noise, maxit=100)
    xr, fr = B.ie_minimize(Imea, Amat, Ie0, noise)
    print(f"{noise}: newton f={f:.15e} ({its} it), L-BFGS-B f={fr:.15e}, x={x}")
    assert abs(f - fr) <= 1e-9 * abs(fr)
    _, g, _ = obj(Imea, Amat, x)
    gs = np.abs(obj(Imea, Amat, Ie0)[1]).max()
    assert np.all(np.abs(g[x > 0]) <= 1e-7 * gs)
    assert np.all(g[x == 0] >= -1e-7 * gs)
    assert np.sum(x == 0) >= 1


@pytest.mark.parametrize("noise", ["poisson", "gaussian"])
def test_ie_newton_ill_conditioned_descends(noise):
    """E = 20 (condition number ~1e18): every accepted iterate lowers f, the solver ends
    within 1e-3 of L-BFGS-B's objective or below it, and a planted ascent step (the sign
    of d flipped) is rejected by the sufficient-decrease test."""
    Amat, Imea, Ie0 = _ie_problem(12, noise)
    obj = B.gauss_l_i if noise == "gaussian" else B.poiss_l_i
    tr = []
    x, f, its = B.ie_newton(Imea, Amat, Ie0, noise, maxit=200, trace=tr)
    assert np.all(np.diff(tr) < 0)
    _, fr = B.ie_minimize(Imea, Amat, Ie0, noise)
    print(f"{noise}: newton f={f:.12e} after {its} it, L-BFGS-B f={fr:.12e}")
    assert f <= fr + 1e-3 * abs(fr)
    f0, g0, _ = obj(Imea, Amat, Ie0)
    up = np.maximum(Ie0 + 1e-2 * g0 / np.abs(g0).max() * Ie0.max(), 0)
    assert obj(Imea, Amat, up)[0] > f0 - 1e-4 * (g0 @ (Ie0 - up))
