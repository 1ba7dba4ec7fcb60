"""GPU parity at the benchmark configurations' own sizes.

* 1024^2 fan slice (360 views, 1024 bins, d = 4096): forward and adjoint against the
  reference's own cpuPrj (prj/cpuPrj.c:679-806, compiled into oracle/_ref) and against
  the FP64 oracle, <= 1e-5 relative L2 each, plus the adjoint identity (<= 1e-6).
* 512^2 fan beam-hardening problem (360 views x 729 bins, d = 2000, E = 20 spline
  bases, Poisson counts at I0 = 1e5): f, grad_alpha, the Hessian ray weight
  (bhc/poissLAlpha.m), f, grad_Ie and the E x E Hessian (bhc/poissLI.m) against the FP64
  restatement; central differences (h = 1e-3) of the device NLL along 3 alpha
  directions and for all 20 Ie coordinates.
* 64 ragged slices in one plan (n in {128..1024}, 90..720 views, mixed parallel / fan):
  every slice with n <= 512 against the FP64 oracle at the kernels' angles (<= 1e-5
  relative L2, forward and adjoint), every slice by the adjoint identity (<= 1e-6 of
  <|x|, A^T|y|>, the sum of the magnitudes of its terms).
Errors against FP64 are printed per result."""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_projector import ANGLE_MODE

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32).ravel()).cuda()


def test_fan1024_against_reference_and_fp64(oracle_mod):
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.workloads import FAN1024
    geo = FAN1024
    n, W, npj = geo[:3]
    P = Projector([geo])
    img = ellipse_phantom(n, 2)
    s = P.forward(_t(img)).cpu().numpy().reshape(npj, W)
    y = gaussian((npj, W), 22)
    b = P.adjoint(_t(y)).cpu().numpy().reshape(n, n)
    bs = P.adjoint(_t(s)).cpu().numpy().reshape(n, n)
    cf = oracle_mod.conf(*geo, angle_mode=ANGLE_MODE[P.modes()[0]])
    s64, _ = oracle_mod.forward(cf, img)
    b64, _ = oracle_mod.adjoint(cf, y)
    mag, _ = oracle_mod.adjoint(cf, np.abs(y))
    e_f, e_b = rel_l2(s, s64), np.linalg.norm(b - b64) / np.linalg.norm(mag)
    print(f"fan1024 vs FP64: fwd {e_f:.2e}, adj(noise) {e_b:.2e} of |A^T||y|")
    assert e_f <= 1e-5 and e_b <= 1e-5
    if oracle_mod.ref_available():
        sr = oracle_mod.ref_run(*geo, img, oracle_mod.FWD_BIT)
        br = oracle_mod.ref_run(*geo, s, oracle_mod.BWD_BIT)
        e_fr, e_br = rel_l2(s, sr), rel_l2(bs, br)
        print(f"fan1024 vs reference cpuPrj: fwd {e_fr:.2e}, adj(A x) {e_br:.2e}")
        assert e_fr <= 1e-5 and e_br <= 1e-5
        assert np.all(sr[:, 0] == 0) and np.all(s[:, 0] == 0)
        assert np.all(bs[0, :] == 0) and np.all(bs[:, 0] == 0)
    x = gaussian((n, n), 21)
    lhs = torch.dot(P.forward(_t(x)).double(), _t(y).double()).item()
    rhs = torch.dot(_t(x).double(), P.adjoint(_t(y)).double()).item()
    print(f"fan1024 adjoint identity {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


# ------------------------------------------------- 512^2 beam-hardening problem ---

E = 20


@pytest.fixture(scope="module")
def bhc512(oracle_mod):
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.workloads import FAN512
    geo = FAN512
    n = geo[0]
    knots = B.spline_knots(B.kappa_logspace(E, 0.03, 30.0))
    x = np.arange(E)
    Ie = np.exp(-0.5 * ((x - 7.5) / 3.0) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    truth = rasterise(ellipse_params(3, n, amp_lo=0.0, amp_hi=1.0), n)
    P = Projector([geo])
    cf = oracle_mod.conf(*geo, angle_mode=ANGLE_MODE[P.modes()[0]])
    raw, _ = oracle_mod.forward(cf, truth)
    Ts = 3.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw.ravel(), Ie)
    counts = np.random.default_rng(3).poisson(1e5 * BL)
    Imea = counts / counts.max()
    alpha = truth * 0.8 + 0.05
    return dict(geo=geo, P=P, cf=cf, knots=knots, Ie=Ie, Ts=Ts, Imea=Imea, alpha=alpha,
                Imea_t=_t(Imea), alpha_t=_t(alpha))


def test_bhc512_against_fp64(bhc512, oracle_mod):
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.bhc import poissLAlpha, poissLI
    Q = bhc512
    geo, cf, Ts = Q["geo"], Q["cf"], Q["Ts"]
    n, W, npj = geo[:3]
    f, g, hw = poissLAlpha(Q["Imea_t"], Q["Ie"], Q["alpha_t"], Q["P"], Ts, Q["knots"],
                           want_hess=True)
    Phi = lambda a: Ts * oracle_mod.forward(cf, a.reshape(n, n))[0].ravel()  # noqa: E731
    Phit = lambda y: Ts * oracle_mod.adjoint(cf, y.reshape(npj, W))[0].ravel()  # noqa: E731
    pol = lambda k, s, I: B.b1_iout_stable(k, s, I)  # noqa: E731
    f64, g64, hw64 = B.poiss_l_alpha(Q["Imea"], Q["Ie"], Q["alpha"].ravel(), Phi, Phit, pol,
                                     Q["knots"])
    ef, eg = abs(f.item() - f64) / abs(f64), rel_l2(g.cpu().numpy(), g64)
    eh = rel_l2(hw.cpu().numpy(), hw64)
    print(f"bhc512 poissLAlpha vs FP64: f {ef:.2e}, grad {eg:.2e}, Hessian weight {eh:.2e}")
    assert ef <= 1e-5 and eg <= 1e-5 and eh <= 1e-4
    fi, gi, H = poissLI(Q["Imea_t"], Q["Ie"], Q["P"], Ts, Q["knots"], alpha=Q["alpha_t"])
    A, _, _ = B.b1_iout_stable(Q["knots"], Phi(Q["alpha"].ravel()), None)
    fi64, gi64, H64 = B.poiss_l_i(Q["Imea"], A, Q["Ie"])
    efi, egi = abs(fi.item() - fi64) / abs(fi64), rel_l2(gi.cpu().numpy(), gi64)
    eH = rel_l2(H.cpu().numpy(), H64)
    print(f"bhc512 poissLI vs FP64: f {efi:.2e}, grad {egi:.2e}, Hessian {eH:.2e}")
    assert efi <= 1e-5 and egi <= 1e-5 and eH <= 1e-5
    Ir = A @ Q["Ie"]
    assert gi.cpu().numpy() @ Q["Ie"] == pytest.approx(np.sum(Ir - Q["Imea"]), rel=1e-5)


def test_bhc512_finite_differences(bhc512):
    """Central differences (h = 1e-3) of the device NLL along 3 seeded alpha directions
    against <grad_alpha, v>, and for every Ie coordinate (Richardson-extrapolated over h
    and h/2) against grad_Ie."""
    import torch
    from imgrecsrc_amd.bhc import poissLAlpha, poissLI
    Q = bhc512
    P, Ts, kn, Ie = Q["P"], Q["Ts"], Q["knots"], Q["Ie"]
    n = Q["geo"][0]
    _, g = poissLAlpha(Q["Imea_t"], Ie, Q["alpha_t"], P, Ts, kn)
    rng = np.random.default_rng(17)
    h = 1e-3
    for k in range(3):
        v = _t(rng.standard_normal(n * n))
        fp = poissLAlpha(Q["Imea_t"], Ie, Q["alpha_t"] + h * v, P, Ts, kn, want_grad=False).item()
        fm = poissLAlpha(Q["Imea_t"], Ie, Q["alpha_t"] - h * v, P, Ts, kn, want_grad=False).item()
        fd, gv = (fp - fm) / (2 * h), torch.dot(g.double(), v.double()).item()
        print(f"bhc512 alpha direction {k}: FD {fd:.6e} vs <g,v> {gv:.6e} "
              f"({abs(fd - gv) / abs(gv):.2e})")
        assert abs(fd - gv) <= 2e-3 * abs(gv)
    _, gi, _ = poissLI(Q["Imea_t"], Ie, P, Ts, kn, alpha=Q["alpha_t"])
    gi = gi.cpu().numpy()

    def fie(e):
        return poissLI(Q["Imea_t"], Ie + e, P, Ts, kn, alpha=Q["alpha_t"])[0].item()

    worst = 0.0
    for j in range(E):
        e = np.zeros(E)
        e[j] = h
        d1 = (fie(e) - fie(-e)) / (2 * h)
        d2 = (fie(e / 2) - fie(-e / 2)) / h
        rich = (4 * d2 - d1) / 3
        worst = max(worst, abs(rich - gi[j]) / np.abs(gi).max())
        assert abs(rich - gi[j]) <= 1e-5 * np.abs(gi).max(), (j, rich, gi[j])
    print(f"bhc512 Ie finite differences: worst {worst:.2e} of max |g|")


# ------------------------------------------------------------ 64 ragged slices ---

def test_ragged64_against_fp64(oracle_mod):
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.workloads import ragged_geoms
    geoms = ragged_geoms(seed=4)
    B = Projector(geoms)
    B.autotune(reps=1)
    modes = B.modes()
    imgs = [ellipse_phantom(g[0], 100 + i) for i, g in enumerate(geoms)]
    ys = [gaussian((g[2], g[1]), 300 + i) for i, g in enumerate(geoms)]
    xs = [gaussian((g[0], g[0]), 500 + i) for i, g in enumerate(geoms)]
    sb = [t.cpu().numpy() for t in B.unpack_sinos(B.forward(B.pack_images([_t(a) for a in imgs])))]
    yb = B.pack_sinos([_t(a) for a in ys])
    ib = [t.cpu().numpy() for t in B.unpack_images(B.adjoint(yb))]
    Ax = B.unpack_sinos(B.forward(B.pack_images([_t(a) for a in xs])))
    # A^T|y|: <|x|, A^T|y|> bounds the magnitude of the terms of <Ax, y> (Gaussian data
    # makes the sum itself cancel)
    ab = [t.cpu().numpy() for t in B.unpack_images(B.adjoint(B.pack_sinos([_t(np.abs(a)) for a in ys])))]
    checked = 0
    for i, g in enumerate(geoms):
        n, W, npj = g[:3]
        # adjoint identity on every slice
        lhs = float(np.dot(Ax[i].cpu().numpy().ravel().astype(np.float64), ys[i].ravel()))
        rhs = float(np.dot(xs[i].ravel(), ib[i].ravel().astype(np.float64)))
        scale = float(np.dot(np.abs(xs[i]).ravel(), ab[i].ravel().astype(np.float64)))
        assert abs(lhs - rhs) <= 1e-6 * scale, (i, g, lhs, rhs, scale)
        if n > 512:
            continue
        cf = oracle_mod.conf(*g, angle_mode=ANGLE_MODE[modes[i]])
        M = oracle_mod.Matrix(cf)
        e_f = rel_l2(sb[i], M.forward(imgs[i])[0])
        b64 = M.adjoint(ys[i])[0]
        mag = M.adjoint(np.abs(ys[i]))[0]
        e_b = np.linalg.norm(ib[i].reshape(n, n) - b64) / np.linalg.norm(mag)
        print(f"slice {i} {g} group {modes[i]}: fwd {e_f:.2e}, adj {e_b:.2e} vs FP64, "
              f"identity {abs(lhs - rhs) / scale:.1e} of <|x|,A^T|y|>")
        assert e_f <= 1e-5 and e_b <= 1e-5, (i, g)
        checked += 1
        del M
    assert checked == sum(1 for g in geoms if g[0] <= 512)
    torch.cuda.synchronize()
