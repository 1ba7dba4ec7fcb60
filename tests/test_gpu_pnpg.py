"""On-device PNPG (ctprj_pnpg_*) against the FP64 restatement of the same trial-by-trial
method (oracle/pnpg_oracle.py) on a small ragged stack, and HIP-graph replay of captured
trials against eager execution (bit for bit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOMS = [(48, 71, 60, 60, 1.0, 1.0, 200.0), (40, 61, 45, 90, 1.0, 1.0, 0.0),
         (33, 50, 48, 48, 1.0, 1.0, 150.0)]
E = 8


@pytest.fixture(scope="module")
def problem(oracle_mod):
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    knots = B.spline_knots(B.kappa_logspace(E, 0.05, 20.0))
    Ie = np.exp(-0.5 * ((np.arange(E) - 3.0) / 1.5) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    truth, sinos, cfs = [], [], []
    for i, g in enumerate(GEOMS):
        a = rasterise(ellipse_params(60 + i, g[0]), g[0])
        cf = oracle_mod.conf(*g)
        s, _ = oracle_mod.forward(cf, a)
        truth.append(a)
        sinos.append(s.ravel())
        cfs.append(cf)
    raw = np.concatenate(sinos)
    Ts = 2.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw, Ie)
    rng = np.random.default_rng(70)
    counts = rng.poisson(1e5 * BL)
    Imea = counts / counts.max()
    alpha0 = np.concatenate([np.full(g[0] * g[0], 0.05) for g in GEOMS])
    return dict(knots=knots, Ie=Ie, Ts=Ts, Imea=Imea, alpha0=alpha0, cfs=cfs,
                Imea_t=torch.from_numpy(Imea.astype(np.float32)).cuda())


def _oracle_func(oracle_mod, P):
    from oracle import bhc_oracle as B

    def func(x, grad=True):
        offs = np.cumsum([0] + [g[0] * g[0] for g in GEOMS])
        s = np.concatenate([oracle_mod.forward(cf, x[offs[i]:offs[i + 1]].reshape(g[0], g[0]))[0].ravel()
                            for i, (cf, g) in enumerate(zip(P["cfs"], GEOMS))])
        BL, sBL, _ = B.b1_iout_stable(P["knots"], P["Ts"] * s, P["Ie"])
        f = B._nll(BL, P["Imea"])
        if not grad:
            return f, None
        w = P["Ts"] * (1 - P["Imea"] / BL) * sBL
        so = np.cumsum([0] + [g[2] * g[1] for g in GEOMS])
        gr = np.concatenate([oracle_mod.adjoint(cf, w[so[i]:so[i + 1]].reshape(g[2], g[1]))[0].ravel()
                             for i, (cf, g) in enumerate(zip(P["cfs"], GEOMS))])
        return f, gr
    return func


@pytest.mark.parametrize("method", ["PNPG", "NPG"])
def test_pnpg_matches_fp64_restatement(problem, oracle_mod, method):
    import torch
    from oracle.pnpg_oracle import PnpgOracle
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector(GEOMS)
    alpha = torch.from_numpy(P["alpha0"].astype(np.float32)).cuda()
    solver = PNPG(proj, P["Imea_t"], P["knots"], P["Ie"], P["Ts"], alpha)
    solver.set_method(method)
    ref = PnpgOracle(_oracle_func(oracle_mod, P), P["alpha0"], solver.t0, npg=method == "NPG")
    costs = []
    for k in range(12):
        solver.step(alpha)
        st = solver.stats()
        ref.trial()
        costs.append((st["cost"], ref.cost, st["last_accepted"], ref.log[-1][0]))
        assert (st["last_accepted"] == 1) == (ref.log[-1][0] == "accept"), costs
        if np.isfinite(ref.cost):
            assert abs(st["cost"] - ref.cost) <= 1e-5 * abs(ref.cost), costs
    assert st["iterations"] == ref.p and st["iterations"] >= 6
    # the objective decreased along the accepted iterates
    acc = [c[1] for c in costs if c[3] == "accept"]
    assert acc[-1] < acc[0]
    a = alpha.cpu().numpy()
    assert np.linalg.norm(a - ref.alpha) <= 1e-4 * np.linalg.norm(ref.alpha)


@pytest.mark.parametrize("splits", [(1, 1), (3, 2)])
def test_pnpg_graph_replay_matches_eager(problem, splits):
    """Captured trials replayed R times, with eager work between the replays, equal the
    eager trials bit for bit; (3, 2): the projector's split work items and their combine
    launches inside the graph."""
    import torch
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    P = problem
    proj = Projector(GEOMS)
    proj.set_tune(adj_split=splits[0], fwd_split=splits[1])
    a_eager = torch.from_numpy(P["alpha0"].astype(np.float32)).cuda()
    a_graph = a_eager.clone()
    s_eager = PNPG(proj, P["Imea_t"], P["knots"], P["Ie"], P["Ts"], a_eager, t0=50.0)
    s_graph = PNPG(proj, P["Imea_t"], P["knots"], P["Ie"], P["Ts"], a_graph, t0=50.0)
    K, R = 4, 3
    for _ in range(K * R):
        s_eager.step(a_eager)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    with torch.cuda.graph(g, stream=cap):
        for _ in range(K):
            s_graph.step(a_graph)
    for r in range(R):
        g.replay()
        torch.cuda.synchronize()
        # eager work between replays, as a caller would interleave it
        _ = proj.forward(a_graph.clone())
        torch.cuda.synchronize()
    assert s_graph.stats()["trials"] == K * R
    assert torch.equal(a_graph, a_eager)
    assert s_graph.stats()["cost"] == s_eager.stats()["cost"]


@pytest.mark.parametrize("method", ["PNPG", "NPG"])
def test_pnpg_wavelet_prox_matches_fp64_restatement(oracle_mod, method):
    """The default BHC prox ('wvltADMM'): trials against the FP64 restatement with the
    wavelet ADMM of npg/admm.m; decisions identical, objective (NLL + u ||W alpha||_1) to
    1e-5, iterate to 1e-4 of its scale."""
    import torch
    from oracle import bhc_oracle as B
    from oracle.pnpg_oracle import PnpgWaveletOracle
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf
    geo = (64, 91, 72, 72, 1.0, 1.0, 250.0)
    n = geo[0]
    knots = B.spline_knots(B.kappa_logspace(E, 0.05, 20.0))
    Ie = np.exp(-0.5 * ((np.arange(E) - 3.0) / 1.5) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    cf = oracle_mod.conf(*geo)
    truth = rasterise(ellipse_params(81, n), n)
    raw = oracle_mod.forward(cf, truth)[0].ravel()
    Ts = 2.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw, Ie)
    counts = np.random.default_rng(82).poisson(1e5 * BL)
    Imea = counts / counts.max()

    def func(x, grad=True):
        s = oracle_mod.forward(cf, x.reshape(n, n))[0].ravel()
        BLx, sBL, _ = B.b1_iout_stable(knots, Ts * s, Ie)
        f = B._nll(BLx, Imea)
        if not grad:
            return f, None
        w = Ts * (1 - Imea / BLx) * sBL
        return f, oracle_mod.adjoint(cf, w.reshape(geo[2], geo[1]))[0].ravel()

    h = daubcqf(4)[0]
    L = 3
    proj = Projector([geo])
    alpha0 = np.full(n * n, 0.05)
    alpha = torch.from_numpy(alpha0.astype(np.float32)).cuda()
    solver = PNPG(proj, torch.from_numpy(Imea.astype(np.float32)).cuda(), knots, Ie, Ts, alpha)
    u = 2e-3 * solver.t0
    wv = Wavelet(n, n, h, L)
    solver.set_method(method)
    solver.set_wavelet(wv, u)
    ref = PnpgWaveletOracle(func, alpha0, solver.t0, u, h, L, (n, n), npg=method == "NPG")
    log = []
    for k in range(12):
        solver.step(alpha)
        st = solver.stats()
        ref.trial()
        log.append((ref.log[-1][0], st["last_accepted"], st["cost"], ref.cost, st["innerSearch"],
                    ref.innerSearch, st["t"], ref.t))
        assert st["t"] == pytest.approx(ref.t, rel=1e-12), log
        assert (st["last_accepted"] == 1) == (ref.log[-1][0] == "accept"), log
        if np.isfinite(ref.cost):
            assert abs(st["cost"] - ref.cost) <= 1e-5 * abs(ref.cost), log
    print("\n".join(map(str, log)))
    # NPG spends its first iterations growing the step (preSteps)
    assert st["iterations"] == ref.p and st["iterations"] >= (3 if method == "NPG" else 5)
    got = alpha.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref.alpha).max() <= 1e-4 * np.abs(ref.alpha).max()
    assert 0 < st["innerSearch"] <= st["maxInnerItr"]


def test_pnpg_masked_wavelet_prior_matches_restatement(oracle_mod):
    """alpha restricted to a disc support (opt.mask) with the wavelet prior on the
    supports (tomography/ConfigCT.m:85-86): pixels outside stay zero, the trials match
    the restatement on the supports' vectors."""
    import torch
    from oracle import bhc_oracle as B
    from oracle.pnpg_oracle import PnpgWaveletOracle
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf, wvlt_mask
    geo = (64, 91, 72, 72, 1.0, 1.0, 250.0)
    n = geo[0]
    c = np.arange(n) - (n - 1) / 2
    X, Y = np.meshgrid(c, c)
    mask = X ** 2 + Y ** 2 <= 30 ** 2
    knots = B.spline_knots(B.kappa_logspace(E, 0.05, 20.0))
    Ie = np.exp(-0.5 * ((np.arange(E) - 3.0) / 1.5) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    cf = oracle_mod.conf(*geo)
    truth = rasterise(ellipse_params(81, n), n) * mask
    raw = oracle_mod.forward(cf, truth)[0].ravel()
    Ts = 2.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw, Ie)
    counts = np.random.default_rng(83).poisson(1e5 * BL)
    Imea = counts / counts.max()

    def func(x, grad=True):
        s = oracle_mod.forward(cf, x.reshape(n, n))[0].ravel()
        BLx, sBL, _ = B.b1_iout_stable(knots, Ts * s, Ie)
        f = B._nll(BLx, Imea)
        if not grad:
            return f, None
        w = Ts * (1 - Imea / BLx) * sBL
        return f, oracle_mod.adjoint(cf, w.reshape(geo[2], geo[1]))[0].ravel()

    h = daubcqf(4)[0]
    L = 3
    mk = wvlt_mask(mask, h, L)
    proj = Projector([geo])
    alpha0 = np.where(mask, 0.05, 0.0).ravel()
    alpha = torch.from_numpy(alpha0.astype(np.float32)).cuda()
    solver = PNPG(proj, torch.from_numpy(Imea.astype(np.float32)).cuda(), knots, Ie, Ts, alpha)
    u = 2e-3 * solver.t0
    wv = Wavelet(n, n, h, L)
    wv.set_masks(mask, mk)
    solver.set_mask(mask)
    solver.set_wavelet(wv, u)
    ref = PnpgWaveletOracle(func, alpha0, solver.t0, u, h, L, (n, n), mask=mask.ravel(),
                            coef_mask=mk)
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
    got = alpha.cpu().numpy().astype(np.float64)
    assert not got[~mask.ravel()].any()
    assert st["iterations"] == ref.p and st["iterations"] >= 5
    assert np.abs(got - ref.alpha).max() <= 1e-4 * np.abs(ref.alpha).max()


def test_pnpg_reset_step_estimate_matches_restatement(oracle_mod):
    """Given-up line searches (forced: a failed majorization test gives up at once)
    call reset() (npg/PNPG.m:168-172): theta = 1 and t = min(t, stepSizeInit('hessian'))
    with the Hessian step Ts^2 <A g, hw A g>/||g||^2 at the current iterate
    (npg/Methods.m:108-115, bhc/poissLAlpha.m:15-26).  The device's estimates match the
    FP64 restatement's (1e-5), and so do the decisions and the objective."""
    import torch
    from oracle import bhc_oracle as B
    from oracle.pnpg_oracle import PnpgWaveletOracle
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf
    geo = (64, 91, 72, 72, 1.0, 1.0, 250.0)
    n = geo[0]
    knots = B.spline_knots(B.kappa_logspace(E, 0.05, 20.0))
    Ie = np.exp(-0.5 * ((np.arange(E) - 3.0) / 1.5) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    cf = oracle_mod.conf(*geo)
    truth = rasterise(ellipse_params(81, n), n)
    raw = oracle_mod.forward(cf, truth)[0].ravel()
    Ts = 2.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw, Ie)
    counts = np.random.default_rng(82).poisson(1e5 * BL)
    Imea = counts / counts.max()
    A = lambda x: oracle_mod.forward(cf, x.reshape(n, n))[0].ravel()  # noqa: E731
    At = lambda y: oracle_mod.adjoint(cf, y.reshape(geo[2], geo[1]))[0].ravel()  # noqa: E731

    def func(x, grad=True):
        BLx, sBL, _ = B.b1_iout_stable(knots, Ts * A(x), Ie)
        f = B._nll(BLx, Imea)
        if not grad:
            return f, None
        return f, At(Ts * (1 - Imea / BLx) * sBL)

    def hess_step(x):
        BLx, sBL, ssBL = B.b1_iout_stable(knots, Ts * A(x), Ie)
        g = At(Ts * (1 - Imea / BLx) * sBL)
        hw = (sBL / BLx) ** 2 * Imea + ssBL * (1 - Imea / BLx)
        Ag = Ts * A(g)
        return float(hw @ (Ag * Ag) / (g @ g))

    h = daubcqf(4)[0]
    L = 3
    proj = Projector([geo])
    alpha0 = np.full(n * n, 0.05)
    alpha = torch.from_numpy(alpha0.astype(np.float32)).cuda()
    solver = PNPG(proj, torch.from_numpy(Imea.astype(np.float32)).cuda(), knots, Ie, Ts, alpha)
    t0 = 0.3 * solver.t0  # steps too long: the test fails early and often
    solver = PNPG(proj, torch.from_numpy(Imea.astype(np.float32)).cuda(), knots, Ie, Ts, alpha,
                  t0=t0)
    u = 2e-3 * solver.t0
    wv = Wavelet(n, n, h, L)
    solver.set_wavelet(wv, u)
    solver.set_linesearch(0)
    ref = PnpgWaveletOracle(func, alpha0, t0, u, h, L, (n, n), hess_step=hess_step)
    ref.max_ls = 0
    log = []
    for k in range(14):
        solver.step(alpha)
        st = solver.stats()
        ref.trial()
        log.append((ref.log[-1][0], st["last_accepted"], st["cost"], ref.cost, st["t"], ref.t,
                    st["resets"], ref.resets, st["t_reset"], ref.t_reset))
        assert st["resets"] == ref.resets, log
        if ref.resets:
            assert st["t_reset"] == pytest.approx(ref.t_reset, rel=1e-5), log
        assert st["t"] == pytest.approx(ref.t, rel=1e-5), log
        assert (st["last_accepted"] == 1) == (ref.log[-1][0] == "accept"), log
        if np.isfinite(ref.cost):
            assert abs(st["cost"] - ref.cost) <= 1e-5 * abs(ref.cost), log
    print("\n".join(map(str, log)))
    assert ref.resets >= 1


@pytest.mark.parametrize("prior", ["nonneg", "wavelet"])
def test_snapshot_resume_bit_for_bit(oracle_mod, prior, tmp_path):
    """A run saved after 5 trials (PNPG.save: solver state + alpha) and resumed in a new
    solver continues exactly as the original: 6 more trials give the same alpha and
    statistics bit for bit (bhc/doBHC.m:85-91 saves and warm-starts between levels)."""
    import torch
    from oracle import bhc_oracle as B
    from imgrecsrc_amd._lib import CtprjError
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf
    geo = (64, 91, 72, 72, 1.0, 1.0, 250.0)
    n = geo[0]
    knots = B.spline_knots(B.kappa_logspace(E, 0.05, 20.0))
    Ie = np.exp(-0.5 * ((np.arange(E) - 3.0) / 1.5) ** 2) + 0.05
    Ie = Ie / (((knots[2:] - knots[:-2]) / 2) @ Ie)
    cf = oracle_mod.conf(*geo)
    raw = oracle_mod.forward(cf, rasterise(ellipse_params(91, n), n))[0].ravel()
    Ts = 2.0 / raw.max()
    BL, _, _ = B.b1_iout_stable(knots, Ts * raw, Ie)
    counts = np.random.default_rng(92).poisson(1e5 * BL)
    Imea = torch.from_numpy((counts / counts.max()).astype(np.float32)).cuda()
    proj = Projector([geo])

    def make(alpha, t0=None):
        s = PNPG(proj, Imea, knots, Ie, Ts, alpha, t0=t0)
        if prior == "wavelet":
            s.set_wavelet(Wavelet(n, n, daubcqf(4)[0], 3), 2e-3 * s.t0)
        return s

    alpha = torch.full((n * n,), 0.05, dtype=torch.float32, device="cuda")
    a = make(alpha)
    for _ in range(5):
        a.step(alpha)
    path = tmp_path / "run.npz"
    a.save(path, alpha)
    for _ in range(6):
        a.step(alpha)
    sa = a.stats()
    alpha_b = torch.zeros_like(alpha)
    b = make(alpha_b.fill_(0.05), t0=a.t0)
    b.load(path, alpha_b)
    for _ in range(6):
        b.step(alpha_b)
    sb = b.stats()
    assert torch.equal(alpha, alpha_b)
    assert sa == sb, (sa, sb)
    assert sa["trials"] == 11
    blob = a.save_state()
    with pytest.raises(CtprjError):
        b.load_state(blob[:-4])
    other = PNPG(Projector([(48, 71, 60, 60, 1.0, 1.0, 200.0)]), torch.ones(60 * 71, device="cuda"),
                 knots, Ie, Ts, torch.full((48 * 48,), 0.05, device="cuda"), t0=1.0)
    with pytest.raises(CtprjError):
        other.load_state(np.zeros(int(blob.size), np.uint8))
