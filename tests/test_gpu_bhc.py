"""GPU parity of the beam-hardening model (bhc/Spline.m b1Iout, bhc/poissLAlpha.m,
bhc/poissLI.m) against the FP64 oracle, plus finite-difference checks of the GPU
gradients and the autograd binding.

FP32 per-ray arithmetic and projections, FP64 reductions.  Bars: BL/sBL/ssBL 1e-5 of the
basis scale; NLL and gradients 1e-5 relative; Hessian 1e-5 relative."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

GEO = (64, 91, 72, 72, 1.0, 1.0, 250.0)  # reduced 512^2/729-bin fan workload
E = 20


@pytest.fixture(scope="module")
def problem(oracle_mod):
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.phantom import rasterise, ellipse_params
    n, W, npj = GEO[0], GEO[1], GEO[2]
    knots = B.spline_knots(B.kappa_logspace(E, 0.03, 30.0))
    x = np.arange(E)
    Ie = np.exp(-0.5 * ((x - 7.5) / 3.0) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    alpha_true = rasterise(ellipse_params(3, n, amp_lo=0.0, amp_hi=1.0), n)
    cf = oracle_mod.conf(*GEO)
    raw, _ = oracle_mod.forward(cf, alpha_true)
    Ts = 3.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw.ravel(), Ie)
    rng = np.random.default_rng(33)
    counts = rng.poisson(1e5 * BL)
    Imea = counts / counts.max()
    alpha = alpha_true * 0.8 + 0.05  # evaluation point away from the truth
    return dict(knots=knots, Ie=Ie, Ts=Ts, Imea=Imea, alpha=alpha, cf=cf,
                Imea_t=torch.from_numpy(Imea.astype(np.float32)).cuda(),
                alpha_t=torch.from_numpy(alpha.astype(np.float32).ravel()).cuda())


def test_b1_iout_device(problem):
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import Spline
    s = np.concatenate([[0.0, 1e-7, 1e-4], np.logspace(-3, 1.3, 500), [-0.02]]).astype(np.float32)
    sp = Spline("b1", problem["knots"])
    st = torch.from_numpy(s).cuda()
    BL, sBL, ssBL = [t.cpu().numpy() for t in sp.polyIout(st, problem["Ie"])]
    r = B.b1_iout_stable(problem["knots"], s.astype(np.float64), problem["Ie"])
    for got, ref in zip((BL, sBL, ssBL), r):
        assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref[0]) + 1e-5 * np.abs(ref))
    A = sp.polyIout(st, None).cpu().numpy()
    Ar, _, _ = B.b1_iout_stable(problem["knots"], s.astype(np.float64), None)
    assert np.all(np.abs(A - Ar) <= 1e-5 * np.abs(Ar[0])[None, :] + 1e-5 * np.abs(Ar))
    # the reference closed form + Taylor branch agrees within its own accuracy target
    Rm = B.b1_iout(problem["knots"], s.astype(np.float64), problem["Ie"])
    assert np.all(np.abs(BL - Rm[0]) <= 3e-6 * Rm[0][0] + 1e-5 * np.abs(Rm[0]))


def _oracle_alpha(oracle_mod, P):
    from oracle import bhc_oracle as B
    cf, Ts = P["cf"], P["Ts"]
    Phi = lambda a: Ts * oracle_mod.forward(cf, a.reshape(GEO[0], GEO[0]))[0].ravel()  # noqa: E731
    Phit = lambda y: Ts * oracle_mod.adjoint(cf, y.reshape(GEO[2], GEO[1]))[0].ravel()  # noqa: E731
    pol = lambda k, s, I: B.b1_iout_stable(k, s, I)  # noqa: E731
    return B.poiss_l_alpha(P["Imea"], P["Ie"], P["alpha"].ravel(), Phi, Phit, pol, P["knots"])


def test_poiss_l_alpha(problem, oracle_mod):
    from imgrecsrc_amd.bhc import poissLAlpha
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    f, g, hw = poissLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"], proj, P["Ts"], P["knots"],
                           want_hess=True)
    f64, g64, hw64 = _oracle_alpha(oracle_mod, P)
    ef, eg = abs(f.item() - f64) / abs(f64), rel_l2(g.cpu().numpy(), g64)
    print(f"poissLAlpha: f rel {ef:.2e}, grad rel L2 {eg:.2e}")
    assert ef <= 1e-5 and eg <= 1e-5
    assert rel_l2(hw.cpu().numpy(), hw64) <= 1e-4


def test_poiss_l_i(problem, oracle_mod):
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import poissLI
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    f, g, H = poissLI(P["Imea_t"], P["Ie"], proj, P["Ts"], P["knots"], alpha=P["alpha_t"])
    raw, _ = oracle_mod.forward(P["cf"], P["alpha"])
    A, _, _ = B.b1_iout_stable(P["knots"], P["Ts"] * raw.ravel(), None)
    f64, g64, H64 = B.poiss_l_i(P["Imea"], A, P["Ie"])
    ef = abs(f.item() - f64) / abs(f64)
    eg, eh = rel_l2(g.cpu().numpy(), g64), rel_l2(H.cpu().numpy(), H64)
    print(f"poissLI: f rel {ef:.2e}, g rel {eg:.2e}, H rel {eh:.2e}")
    assert ef <= 1e-5 and eg <= 1e-5 and eh <= 1e-5
    Hn = H.cpu().numpy()
    assert np.array_equal(Hn, Hn.T)
    # g . Ie = sum(Ir - Imea)
    Ir = A @ P["Ie"]
    assert g.cpu().numpy() @ P["Ie"] == pytest.approx(np.sum(Ir - P["Imea"]), rel=1e-5, abs=1e-6)


def test_gradients_finite_differences(problem):
    """Central differences (h = 1e-3) of the device NLL: along alpha directions against
    the device alpha-gradient, and per Ie coordinate against the device Ie-gradient."""
    import torch
    from imgrecsrc_amd.bhc import poissLAlpha, poissLI
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    f, g = poissLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"], proj, P["Ts"], P["knots"])
    rng = np.random.default_rng(7)
    h = 1e-3
    for _ in range(3):
        v = torch.from_numpy(rng.standard_normal(GEO[0] ** 2).astype(np.float32)).cuda()
        fp = poissLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"] + h * v, proj, P["Ts"], P["knots"],
                         want_grad=False).item()
        fm = poissLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"] - h * v, proj, P["Ts"], P["knots"],
                         want_grad=False).item()
        fd, gv = (fp - fm) / (2 * h), torch.dot(g.double(), v.double()).item()
        assert abs(fd - gv) <= 2e-3 * abs(gv), (fd, gv)
    _, gi, _ = poissLI(P["Imea_t"], P["Ie"], proj, P["Ts"], P["knots"], alpha=P["alpha_t"])
    gi = gi.cpu().numpy()

    def fie(e):
        return poissLI(P["Imea_t"], P["Ie"] + e, proj, P["Ts"], P["knots"], alpha=P["alpha_t"])[0].item()

    for j in range(0, E, 3):
        e = np.zeros(E)
        e[j] = h
        d1 = (fie(e) - fie(-e)) / (2 * h)
        d2 = (fie(e / 2) - fie(-e / 2)) / h
        rich = (4 * d2 - d1) / 3  # Richardson: O(h^2) truncation -> O(h^4)
        assert abs(rich - gi[j]) <= 1e-5 * np.abs(gi).max(), (j, rich, gi[j])


def test_autograd_binding(problem):
    import torch
    from imgrecsrc_amd.autograd import backproject, poisson_nll, project
    from imgrecsrc_amd.bhc import poissLAlpha, poissLI
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    a = P["alpha_t"].clone().requires_grad_(True)
    ie = torch.tensor(P["Ie"], dtype=torch.float64, requires_grad=True)
    f = poisson_nll(a, ie, proj, P["Imea_t"], P["knots"], P["Ts"])
    f.backward()
    f_ref, g_ref = poissLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"], proj, P["Ts"], P["knots"])
    _, gi_ref, _ = poissLI(P["Imea_t"], P["Ie"], proj, P["Ts"], P["knots"], alpha=P["alpha_t"])
    assert f.item() == f_ref.item()
    assert torch.equal(a.grad, g_ref)
    assert torch.allclose(ie.grad, gi_ref.cpu(), rtol=1e-12, atol=0)
    # projection autograd: d<A x, y>/dx = A^T y
    x = torch.randn(GEO[0] ** 2, device="cuda", requires_grad=True)
    y = torch.randn(GEO[1] * GEO[2], device="cuda")
    (project(x, proj) * y).sum().backward()
    assert torch.equal(x.grad, proj.adjoint(y))
    s = torch.randn(GEO[1] * GEO[2], device="cuda", requires_grad=True)
    (backproject(s, proj) * x.detach()).sum().backward()
    assert torch.equal(s.grad, proj.forward(x.detach()))
    # 2-D inputs: an n x n image and an np x prjWidth sinogram get gradients of their shape
    x2 = x.detach().view(GEO[0], GEO[0]).clone().requires_grad_(True)
    (project(x2, proj) * y).sum().backward()
    assert x2.grad.shape == x2.shape and torch.equal(x2.grad.reshape(-1), proj.adjoint(y))
    s2 = s.detach().view(GEO[2], GEO[1]).clone().requires_grad_(True)
    (backproject(s2, proj) * x.detach()).sum().backward()
    assert s2.grad.shape == s2.shape and torch.equal(s2.grad.reshape(-1), proj.forward(x.detach()))
    a2 = P["alpha_t"].view(GEO[0], GEO[0]).clone().requires_grad_(True)
    poisson_nll(a2, ie.detach(), proj, P["Imea_t"], P["knots"], P["Ts"]).backward()
    assert a2.grad.shape == a2.shape and torch.equal(a2.grad.reshape(-1), g_ref)


def test_out_buffer_checked(problem):
    """A caller's output buffer goes through the input checks plus contiguity: a short,
    wrongly typed or strided `out` raises instead of being written past its end."""
    import torch
    from imgrecsrc_amd.projector import Projector
    proj = Projector([GEO])
    x = problem["alpha_t"]
    y = proj.forward(x)
    bad = [torch.empty(proj.sino_elems - 1, device="cuda"),
           torch.empty(proj.sino_elems, device="cuda", dtype=torch.float64),
           torch.empty(2 * proj.sino_elems, device="cuda")[::2],
           torch.empty(proj.sino_elems)]
    for out in bad:
        with pytest.raises((ValueError, TypeError)):
            proj.forward(x, out=out)
    for out in [torch.empty(proj.img_elems + 7, device="cuda"),
                torch.empty(2 * proj.img_elems, device="cuda")[::2]]:
        with pytest.raises((ValueError, TypeError)):
            proj.adjoint(y, out=out)
        with pytest.raises((ValueError, TypeError)):
            proj.fbp(y, out=out)
    good = torch.full((proj.sino_elems,), 7.0, device="cuda")
    assert proj.forward(x, out=good) is good and torch.equal(good, y)


def test_zero_counts(problem):
    """Imea = 0 rays contribute BL to the NLL (limit of Imea log(BL/Imea)) and stay finite."""
    from imgrecsrc_amd.bhc import poissLAlpha
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    im = P["Imea_t"].clone()
    im[::5] = 0
    f, g = poissLAlpha(im, P["Ie"], P["alpha_t"], proj, P["Ts"], P["knots"])
    assert np.isfinite(f.item()) and bool(g.isfinite().all())


# ---------------------------------------------------------------- model variants ---
# Spectrum bases of bhc/Spline.m ('b0' boxes on E+1 edges, 'dis' nodes) and the Gaussian
# noise model of bhc/gaussLAlpha.m / bhc/gaussLI.m, against the FP64 restatements.

KN0 = np.logspace(np.log10(0.03), np.log10(30.0), E + 1)


def _basis(P, basis):
    from oracle import bhc_oracle as B
    if basis == "b0":
        return KN0, (lambda k, s, I: B.b0_iout_stable(k, s, I))
    if basis == "dis":
        return P["knots"], (lambda k, s, I: B.dis_iout(k, s, I))
    return P["knots"], (lambda k, s, I: B.b1_iout_stable(k, s, I))


@pytest.mark.parametrize("basis", ["b0", "dis"])
def test_poly_iout_bases_device(problem, basis):
    import torch
    from imgrecsrc_amd.bhc import Spline
    s = np.concatenate([[0.0, 1e-7, 1e-4], np.logspace(-3, 1.3, 500), [-0.02]]).astype(np.float32)
    knots, pol = _basis(problem, basis)
    st = torch.from_numpy(s).cuda()
    sp = Spline(basis, knots)
    got = [t.cpu().numpy() for t in sp.polyIout(st, problem["Ie"])]
    ref = pol(knots, s.astype(np.float64), problem["Ie"])
    for g, r in zip(got, ref):
        assert np.all(np.abs(g - r) <= 1e-5 * np.abs(r[0]) + 1e-5 * np.abs(r))
    A = sp.polyIout(st, None).cpu().numpy()
    Ar = pol(knots, s.astype(np.float64), None)[0]
    assert np.all(np.abs(A - Ar) <= 1e-5 * np.abs(Ar[0])[None, :] + 1e-5 * np.abs(Ar))


@pytest.mark.parametrize("basis,noise", [("b1", "gaussian"), ("b0", "poisson"),
                                         ("dis", "poisson"), ("dis", "gaussian")])
def test_alpha_objective_variants(problem, oracle_mod, basis, noise):
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import gaussLAlpha, poissLAlpha
    from imgrecsrc_amd.projector import Projector
    P = problem
    knots, pol = _basis(P, basis)
    proj = Projector([GEO])
    fn = gaussLAlpha if noise == "gaussian" else poissLAlpha
    f, g, hw = fn(P["Imea_t"], P["Ie"], P["alpha_t"], proj, P["Ts"], knots, want_hess=True,
                  basis=basis)
    # without Hessian weights the ray pass runs in the forward projector's store
    f2, g2 = fn(P["Imea_t"], P["Ie"], P["alpha_t"], proj, P["Ts"], knots, basis=basis)
    cf, Ts = P["cf"], P["Ts"]
    Phi = lambda a: Ts * oracle_mod.forward(cf, a.reshape(GEO[0], GEO[0]))[0].ravel()  # noqa: E731
    Phit = lambda y: Ts * oracle_mod.adjoint(cf, y.reshape(GEO[2], GEO[1]))[0].ravel()  # noqa: E731
    ofn = B.gauss_l_alpha if noise == "gaussian" else B.poiss_l_alpha
    f64, g64, hw64 = ofn(P["Imea"], P["Ie"], P["alpha"].ravel(), Phi, Phit, pol, knots)
    for ff, gg, how in ((f, g, "ray pass"), (f2, g2, "fused store")):
        ef, eg = abs(ff.item() - f64) / abs(f64), rel_l2(gg.cpu().numpy(), g64)
        print(f"{noise}/{basis} alpha objective ({how}): f rel {ef:.2e}, grad rel L2 {eg:.2e}")
        assert ef <= 1e-5 and eg <= 1e-5
    assert rel_l2(hw.cpu().numpy(), hw64) <= 1e-4


@pytest.mark.parametrize("basis,noise", [("b1", "gaussian"), ("b0", "poisson"),
                                         ("dis", "gaussian")])
def test_ie_objective_variants(problem, oracle_mod, basis, noise):
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import gaussLI, poissLI
    from imgrecsrc_amd.projector import Projector
    P = problem
    knots, pol = _basis(P, basis)
    proj = Projector([GEO])
    fn = gaussLI if noise == "gaussian" else poissLI
    f, g, H = fn(P["Imea_t"], P["Ie"], proj, P["Ts"], knots, alpha=P["alpha_t"], basis=basis)
    raw, _ = oracle_mod.forward(P["cf"], P["alpha"])
    A = pol(knots, P["Ts"] * raw.ravel(), None)[0]
    ofn = B.gauss_l_i if noise == "gaussian" else B.poiss_l_i
    f64, g64, H64 = ofn(P["Imea"], A, P["Ie"])
    ef = abs(f.item() - f64) / abs(f64)
    eg, eh = rel_l2(g.cpu().numpy(), g64), rel_l2(H.cpu().numpy(), H64)
    print(f"{noise}/{basis} Ie objective: f rel {ef:.2e}, g rel {eg:.2e}, H rel {eh:.2e}")
    assert ef <= 1e-5 and eg <= 1e-5 and eh <= 1e-5


def test_gaussian_gradient_finite_differences(problem):
    """Central differences of the device Gaussian NLL along alpha directions."""
    import torch
    from imgrecsrc_amd.bhc import gaussLAlpha
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    f, g = gaussLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"], proj, P["Ts"], P["knots"], basis="dis")
    rng = np.random.default_rng(9)
    h = 1e-3
    for _ in range(3):
        v = torch.from_numpy(rng.standard_normal(GEO[0] ** 2).astype(np.float32)).cuda()
        fp = gaussLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"] + h * v, proj, P["Ts"], P["knots"],
                         want_grad=False, basis="dis").item()
        fm = gaussLAlpha(P["Imea_t"], P["Ie"], P["alpha_t"] - h * v, proj, P["Ts"], P["knots"],
                         want_grad=False, basis="dis").item()
        fd, gv = (fp - fm) / (2 * h), torch.dot(g.double(), v.double()).item()
        assert abs(fd - gv) <= 2e-3 * abs(gv), (fd, gv)


@pytest.mark.parametrize("noise", ["poisson", "gaussian"])
def test_ie_solve_device(problem, oracle_mod, noise):
    """The I-epsilon step on the device (ctprj_ie_solve): on a well-conditioned spectrum
    model (E = 6) it reaches the FP64 minimiser of scipy's bounded L-BFGS-B, with the
    empty bin found; on the realistic one (E = 20) it lowers the FP64 objective."""
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import gaussLI, ie_solve, poissLI
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector([GEO])
    raw, _ = oracle_mod.forward(P["cf"], P["alpha"])
    s = P["Ts"] * raw.ravel()
    obj = B.gauss_l_i if noise == "gaussian" else B.poiss_l_i
    rng = np.random.default_rng(21)
    for Ek in (6, 20):
        kn = B.spline_knots(B.kappa_logspace(Ek, 0.03, 30.0))
        A = B.b1_iout_stable(kn, s, None)[0]
        truth = np.exp(-0.5 * ((np.arange(Ek) - Ek / 2) / (Ek / 6)) ** 2)
        truth[-1] = 0.0
        truth /= ((kn[2:] - kn[:-2]) / 2) @ truth
        Ir = A @ truth
        Imea = (rng.poisson(1e5 * Ir) / 1e5 if noise == "poisson"
                else Ir * np.exp(0.01 * rng.standard_normal(Ir.size)))
        Ie0 = np.full(Ek, 1.0) / ((kn[2:] - kn[:-2]) / 2).sum()
        im = torch.from_numpy(Imea.astype(np.float32)).cuda()
        x, info = ie_solve(im, Ie0, proj, P["Ts"], kn, alpha=P["alpha_t"], noise=noise, maxit=100)
        # the step's passes read the basis matrix it formed once; the gradient call
        # evaluates the basis: the same cost bit for bit at the returned spectrum
        fn = gaussLI if noise == "gaussian" else poissLI
        assert fn(im, x, proj, P["Ts"], kn, alpha=P["alpha_t"])[0].item() == info["cost"]
        f_dev = obj(Imea.astype(np.float32).astype(np.float64), A, x)[0]
        f0 = obj(Imea.astype(np.float32).astype(np.float64), A, Ie0)[0]
        print(f"{noise} E={Ek}: f0 {f0:.6e} -> {f_dev:.9e} ({info})")
        assert info["iterations"] >= 1 and f_dev < f0
        if Ek == 6:
            xr, fr = B.ie_minimize(Imea.astype(np.float32).astype(np.float64), A, Ie0, noise)
            print(f"   L-BFGS-B f {fr:.9e}, x {np.round(xr, 5)}, device x {np.round(x, 5)}")
            assert abs(f_dev - fr) <= 1e-5 * abs(fr)
            assert np.allclose(x, xr, rtol=1e-3, atol=1e-4 * np.abs(xr).max())


# (geometry, tune fields): the symmetry groups (D4, V4, Z2), a folded half-turn parallel
# scan, odd and even detectors, one and several work-groups and waves per forward item
FUSED_CASES = [
    ((64, 91, 72, 72, 1.0, 1.0, 250.0), dict(fwd_seg=4, fwd_split=1)),
    ((64, 91, 72, 72, 1.0, 1.0, 250.0), dict(fwd_seg=2, fwd_split=3)),
    ((48, 70, 60, 120, 1.0, 1.0, 0.0), dict(fwd_seg=1, fwd_split=1)),
    ((48, 70, 60, 120, 1.0, 1.0, 0.0), dict(fwd_seg=8, fwd_split=4)),
    ((50, 73, 70, 70, 0.9, 1.0, 300.0), dict(fwd_seg=2, fwd_split=1)),
    ((40, 61, 45, 45, 1.0, 1.0, 200.0), dict(fwd_seg=4, fwd_split=2)),
    ((33, 52, 34, 34, 1.2, 1.0, 160.0), dict(fwd_seg=16, fwd_split=1)),
]


@pytest.mark.parametrize("noise", ["poisson", "gaussian"])
@pytest.mark.parametrize("case", FUSED_CASES)
def test_fused_store_matches_ray_pass(problem, case, noise):
    """The alpha gradient with the ray pass fused into the forward projector's store
    against the separate ray pass (ctprj_bhc_alpha_rays) on the same plan's projection
    followed by its adjoint: the gradients bit for bit (the same weights w, the same
    adjoint), the NLL up to the FP64 summation order (the fused path's NLL partials are
    summed by the adjoint's first wave)."""
    import torch
    from imgrecsrc_amd.bhc import alpha_rays, gaussLAlpha, poissLAlpha
    from imgrecsrc_amd.phantom import ellipse_phantom
    from imgrecsrc_amd.projector import Projector
    geo, tune = case
    proj = Projector([geo])
    assert proj.modes()[0] > 0, "case meant for the symmetric kernels"
    proj.set_tune(**tune)
    rng = np.random.default_rng(geo[0] + geo[2])
    alpha = torch.from_numpy(0.05 + ellipse_phantom(geo[0], 7).ravel()).cuda()
    Imea = torch.from_numpy(rng.uniform(0.05, 1.0, proj.sino_elems).astype(np.float32)).cuda()
    fn = gaussLAlpha if noise == "gaussian" else poissLAlpha
    Ts = 0.02
    f2, g2 = fn(Imea, problem["Ie"], alpha, proj, Ts, problem["knots"])
    f1, w = alpha_rays(proj, proj.forward(alpha), Imea, problem["Ie"], problem["knots"], Ts,
                       True, noise=noise)
    g1 = proj.adjoint(w)
    assert torch.equal(g1, g2), (geo, tune, noise)
    assert abs(f1.item() - f2.item()) <= 1e-12 * abs(f1.item()), (f1.item(), f2.item())
    # cost only (no adjoint, the NLL partials summed by a launch of their own): the same
    # value bit for bit
    f3 = fn(Imea, problem["Ie"], alpha, proj, Ts, problem["knots"], want_grad=False)
    assert f3.item() == f2.item()


def _ie_case(rays, E_, seed):
    """Synthetic ray values s (Ts = 1) and measurements near the model's for E_ bins."""
    from oracle import bhc_oracle as B
    rng = np.random.default_rng(seed)
    knots = B.spline_knots(B.kappa_logspace(E_, 0.03, 30.0))
    x = np.arange(E_)
    Ie = np.exp(-0.5 * ((x - 0.4 * E_) / (0.15 * E_)) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    s = rng.uniform(0.0, 3.0, rays).astype(np.float32)
    A = B.b1_iout_stable(knots, s.astype(np.float64), None)[0]
    # measurements 5-25 % below the model, so that g = A'(1 - Imea/Ir) has no cancellation
    Imea = (A @ Ie) * rng.uniform(0.75, 0.95, rays)
    return knots, Ie, s, Imea, A


@pytest.mark.parametrize("E_", [6, 11, 20, 26, 32])
def test_ie_pass_block_layouts(E_):
    """The I-epsilon pass at spectrum sizes that give 64, 32, 16, 8 and 4 lanes per
    H block (E = 6, 11, 20, 26, 32), from given ray values."""
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import poissLI
    from imgrecsrc_amd.projector import Projector
    proj = Projector([GEO])
    knots, Ie, s, Imea, A = _ie_case(proj.sino_elems, E_, 40 + E_)
    f, g, H = poissLI(torch.from_numpy(Imea.astype(np.float32)).cuda(), Ie, proj, 1.0, knots,
                      sraw=torch.from_numpy(s).cuda())
    f64, g64, H64 = B.poiss_l_i(Imea.astype(np.float32).astype(np.float64), A, Ie)
    ef = abs(f.item() - f64) / abs(f64)
    eg, eh = rel_l2(g.cpu().numpy(), g64), rel_l2(H.cpu().numpy(), H64)
    print(f"E={E_}: f rel {ef:.2e}, g rel {eg:.2e}, H rel {eh:.2e}")
    assert ef <= 1e-5 and eg <= 1e-5 and eh <= 1e-5
    Hn = H.cpu().numpy()
    assert np.array_equal(Hn, Hn.T)


def test_ie_pass_several_tiles_per_group():
    """More than 4096 tiles of 256 rays (1100 x 1024 rays): the pass's groups take two
    tiles each."""
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import poissLI
    from imgrecsrc_amd.projector import Projector
    geo = (32, 1024, 1100, 1100, 1.0, 1.0, 4000.0)
    proj = Projector([geo])
    assert proj.sino_elems > 4096 * 256
    knots, Ie, s, Imea, A = _ie_case(proj.sino_elems, E, 7)
    f, g, H = poissLI(torch.from_numpy(Imea.astype(np.float32)).cuda(), Ie, proj, 1.0, knots,
                      sraw=torch.from_numpy(s).cuda())
    f64, g64, H64 = B.poiss_l_i(Imea.astype(np.float32).astype(np.float64), A, Ie)
    ef = abs(f.item() - f64) / abs(f64)
    eg, eh = rel_l2(g.cpu().numpy(), g64), rel_l2(H.cpu().numpy(), H64)
    print(f"{proj.sino_elems} rays: f rel {ef:.2e}, g rel {eg:.2e}, H rel {eh:.2e}")
    assert ef <= 1e-5 and eg <= 1e-5 and eh <= 1e-5
