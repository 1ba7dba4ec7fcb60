"""ctprj_bhc_grad_alpha_noalloc (the gradient entry point that allocates nothing: its
reduction partials live in the library's static device array) against
ctprj_bhc_grad_alpha through the C ABI: the same kernels in the same order, so f, the
gradient and the Hessian ray weight must be bit-identical.  Two calls in a row, the
second on a smaller plan (the static array keeps the first call's partials; each call
must write every partial it reads).  A plan beyond the array's size is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(geo, seed):
    import torch
    from imgrecsrc_amd.bhc import Spline, kappa_logspace, normalize_ie, spline_knots
    from imgrecsrc_amd.phantom import ellipse_params, rasterise
    from imgrecsrc_amd.projector import Projector
    P = Projector([geo])
    n = geo[0]
    knots = spline_knots(kappa_logspace(20, 0.03, 30.0))
    Ie = normalize_ie(np.exp(-0.5 * ((np.arange(20) - 7.5) / 3.0) ** 2) + 0.05, knots)
    truth = torch.from_numpy(rasterise(ellipse_params(seed, n), n).astype(np.float32)).cuda()
    raw = P.forward(truth.reshape(-1))
    Ts = float(3.0 / raw.max().item())
    BL = Spline("b1", knots).polyIout(raw * Ts, Ie, want=1)[0]
    g = torch.Generator(device="cuda").manual_seed(seed)
    counts = torch.poisson(1e5 * BL, generator=g)
    Imea = (counts / counts.max()).float().contiguous()
    alpha = (truth.reshape(-1) * 0.8 + 0.05).contiguous()
    return P, knots, Ie, Ts, Imea, alpha


def _call(fn, P, knots, Ie, Ts, Imea, alpha):
    import torch
    from imgrecsrc_amd._lib import check, dbl_array
    from imgrecsrc_amd.projector import _ptr
    kp, _kh = dbl_array(knots)
    ip, _ih = dbl_array(Ie)
    f = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.empty(P.img_elems, dtype=torch.float32, device="cuda")
    hw = torch.empty(P.sino_elems, dtype=torch.float32, device="cuda")
    check(fn(P.handle, _ptr(alpha), _ptr(Imea), kp, len(Ie), ip, ctypes.c_float(Ts), _ptr(f),
             _ptr(g), _ptr(hw), None))
    torch.cuda.synchronize()
    return f.cpu(), g.cpu(), hw.cpu()


def test_noalloc_matches_grad_alpha():
    import torch
    from imgrecsrc_amd._lib import lib
    L = lib()
    for geo, seed in (((256, 367, 180, 360, 1.0, 1.0, 0.0), 11),
                      ((96, 139, 72, 72, 1.0, 1.0, 400.0), 12)):
        prob = _problem(geo, seed)
        ref = _call(L.ctprj_bhc_grad_alpha, *prob)
        got = _call(L.ctprj_bhc_grad_alpha_noalloc, *prob)
        print(f"{geo}: f {ref[0].item():.12e} / {got[0].item():.12e}")
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
        assert np.isfinite(got[0].item())


def test_noalloc_refuses_oversized_plan():
    import torch
    from imgrecsrc_amd._lib import dbl_array, lib
    from imgrecsrc_amd.projector import Projector, _ptr
    # 33 slices of 512 x 1024 rays: 17.3 M rays, beyond the 16.7 M of the static array
    P = Projector([(64, 1024, 512, 512, 1.0, 1.0, 0.0)] * 33)
    alpha = torch.zeros(P.img_elems, dtype=torch.float32, device="cuda")
    Imea = torch.ones(P.sino_elems, dtype=torch.float32, device="cuda")
    f = torch.zeros(1, dtype=torch.float64, device="cuda")
    knots = np.arange(22, dtype=np.float64) + 1.0
    kp, _kh = dbl_array(knots)
    ip, _ih = dbl_array(np.ones(20))
    rc = lib().ctprj_bhc_grad_alpha_noalloc(P.handle, _ptr(alpha), _ptr(Imea), kp, 20, ip,
                                            ctypes.c_float(1.0), _ptr(f), None, None, None)
    assert rc != 0
    assert b"static workspace" in lib().ctprj_last_error()
