"""Seeded random geometries (image size, detector width and pitch, views used and per
turn, effectiveRate, parallel or fan): the device forward and adjoint against the FP64
oracle at the angles the chosen kernels use, on the symmetric and the general kernels,
with the 3- and 4-bin adjoint windows, per-lane and LDS-staged adjoint gathers, and
with work items split over several work-groups, plus the adjoint identity.  Bars: 2e-6 of the
FP64 result's scale per entry, and 1e-6 of sum |x| A^T|y| for <Ax,y> = <x,A^T y>."""
import numpy as np
import pytest

from test_gpu_projector import ANGLE_MODE, _proj, torch_dev  # noqa: F401

pytestmark = pytest.mark.gpu


def _geoms(count=24, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(6, 48))
        ds = float(rng.choice([0.7, 0.85, 1.0, 1.0, 1.25]))
        W = int(np.ceil(n * 1.5 / ds)) + int(rng.integers(0, 6))
        P = int(rng.integers(8, 90))
        npj = int(rng.integers(1, P + 1)) if rng.random() < 0.5 else P
        er = float(rng.choice([1.0, 1.0, 0.8, 1.2]))
        d = 0.0 if rng.random() < 0.5 else float(n * rng.uniform(2.5, 8.0))
        out.append((n, W, npj, P, ds, er, d))
    return out


@pytest.mark.parametrize("geo", _geoms())
def test_random_geometry_against_fp64(torch_dev, oracle_mod, geo):
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    n, W, npj = geo[:3]
    x = ellipse_phantom(n, 5)
    y = gaussian((npj, W), 6)
    xt = torch.from_numpy(x).cuda().reshape(-1)
    yt = torch.from_numpy(y).cuda().reshape(-1)
    # (flags, adjoint / forward work-groups per item)
    for flags, split in ((1, 1), (17, 1), (0, 1), (1, 3), (17, 2), (33, 1), (49, 2)):
        P = _proj(geo, flags)
        if split > 1:
            P.set_tune(adj_split=split, fwd_split=5 - split)
        mode = P.modes()[0]
        cf = oracle_mod.conf(*geo, angle_mode=ANGLE_MODE[mode], support_mode=1)
        f = P.forward(xt).cpu().numpy().reshape(npj, W)
        f64, _ = oracle_mod.forward(cf, x)
        mag, _ = oracle_mod.forward(cf, np.abs(x))
        assert np.abs(f - f64).max() <= 2e-6 * max(1.0, mag.max()), (geo, flags, mode)
        b = P.adjoint(yt).cpu().numpy().reshape(n, n)
        b64, _ = oracle_mod.adjoint(cf, y)
        magb, _ = oracle_mod.adjoint(cf, np.abs(y))
        assert np.abs(b - b64).max() <= 2e-6 * max(1.0, magb.max()), (geo, flags, mode)
        # <Ax, y> = <x, A^T y> relative to the sum of the terms' magnitudes (Gaussian y
        # makes the sums cancel)
        lhs = float(np.dot(f.ravel().astype(np.float64), y.ravel()))
        rhs = float(np.dot(x.ravel(), b.ravel().astype(np.float64)))
        scale = float(np.dot(np.abs(x).ravel(), magb.ravel()))
        assert abs(lhs - rhs) <= 1e-6 * scale, (geo, flags)
