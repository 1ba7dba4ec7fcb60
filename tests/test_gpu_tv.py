"""GPU parity of the TV prox (utils/denoise_bound_mod.m via utils/TV.m) and TV value
(utils/tlv.m) against the FP64 restatement, and PNPG with the TV prox against the FP64
restatement of the same trials."""
import numpy as np
import pytest

from oracle import tv_oracle as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tv", ["iso", "l1"])
@pytest.mark.parametrize("shape,lam,eps", [((32, 32), 0.3, 1e-6), ((17, 40), 1.0, 1e-4),
                                           ((64, 64), 0.05, 1e-5)])
def test_denoise_matches_oracle(tv, shape, lam, eps):
    import torch
    from imgrecsrc_amd.tv import TV
    rng = np.random.default_rng(shape[0] + shape[1])
    x = (rng.standard_normal(shape) + 0.5).astype(np.float32)
    w = TV(*shape, tv=tv)
    xt = torch.from_numpy(x).cuda().reshape(-1)
    out, itr = w.denoise(xt, lam, eps, 500)
    ref, itr_ref = T.denoise(x.astype(np.float64), lam, tv, eps, 500)
    got = out.cpu().numpy().reshape(shape)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{tv} {shape} lam={lam}: iterations {itr} (oracle {itr_ref}), max rel diff {err:.2e}")
    assert itr == itr_ref
    assert err <= 1e-6
    assert w.tlv(xt) == pytest.approx(T.tlv(x.astype(np.float64), tv), rel=1e-12)


def test_pnpg_tv_prox_matches_fp64_restatement(oracle_mod):
    import torch
    from oracle import bhc_oracle as B
    from oracle.pnpg_oracle import PnpgTVOracle
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.tv import TV
    E = 8
    geo = (48, 71, 60, 120, 1.0, 1.0, 0.0)
    n = geo[0]
    knots = B.spline_knots(B.kappa_logspace(E, 0.05, 20.0))
    Ie = np.exp(-0.5 * ((np.arange(E) - 3.0) / 1.5) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    cf = oracle_mod.conf(*geo)
    truth = rasterise(ellipse_params(91, n), n)
    raw = oracle_mod.forward(cf, truth)[0].ravel()
    Ts = 2.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw, Ie)
    counts = np.random.default_rng(92).poisson(1e5 * BL)
    Imea = counts / counts.max()

    def func(x, grad=True):
        s = oracle_mod.forward(cf, x.reshape(n, n))[0].ravel()
        BLx, sBL, _ = B.b1_iout_stable(knots, Ts * s, Ie)
        f = B._nll(BLx, Imea)
        if not grad:
            return f, None
        w = Ts * (1 - Imea / BLx) * sBL
        return f, oracle_mod.adjoint(cf, w.reshape(geo[2], geo[1]))[0].ravel()

    proj = Projector([geo])
    alpha0 = np.full(n * n, 0.05)
    alpha = torch.from_numpy(alpha0.astype(np.float32)).cuda()
    solver = PNPG(proj, torch.from_numpy(Imea.astype(np.float32)).cuda(), knots, Ie, Ts, alpha)
    u = 5e-3 * solver.t0
    tv = TV(n, n, tv="l1")
    solver.set_tv(tv, u)
    ref = PnpgTVOracle(func, alpha0, solver.t0, u, "l1", (n, n))
    log = []
    for k in range(10):
        solver.step(alpha)
        st = solver.stats()
        ref.trial()
        log.append((ref.log[-1][0], st["last_accepted"], st["cost"], ref.cost, st["innerSearch"],
                    ref.innerSearch))
        assert (st["last_accepted"] == 1) == (ref.log[-1][0] == "accept"), log
        if np.isfinite(ref.cost):
            assert abs(st["cost"] - ref.cost) <= 1e-5 * abs(ref.cost), log
    print("\n".join(map(str, log)))
    assert st["iterations"] == ref.p and st["iterations"] >= 5
    got = alpha.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref.alpha).max() <= 1e-4 * np.abs(ref.alpha).max()


def test_masked_denoise_matches_oracle():
    """TV.denoise with opt.mask (utils/TV.m:68-79): the masked vector embedded in a zero
    image, denoised, and read back on the mask."""
    import torch
    from imgrecsrc_amd.tv import TV
    n = 40
    c = np.arange(n) - (n - 1) / 2
    X, Y = np.meshgrid(c, c)
    m = X ** 2 + Y ** 2 <= 16 ** 2
    x = (np.random.default_rng(8).standard_normal((n, n)) + 0.5).astype(np.float32)
    w = TV(n, n, tv="l1")
    w.set_mask(m)
    out, itr = w.denoise(torch.from_numpy(x).cuda().reshape(-1), 0.4, 1e-6, 500)
    ref, itr_ref = T.denoise(np.where(m, x, 0).astype(np.float64), 0.4, "l1", 1e-6, 500)
    got = out.cpu().numpy().reshape(n, n)
    assert itr == itr_ref
    assert not got[~m].any()
    assert np.abs(got[m] - ref[m]).max() <= 1e-6 * np.abs(ref).max()
    assert w.tlv(torch.from_numpy(x).cuda().reshape(-1)) == pytest.approx(
        T.tlv(np.where(m, x, 0).astype(np.float64), "l1"), rel=1e-12)
