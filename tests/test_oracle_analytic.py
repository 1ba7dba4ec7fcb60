"""Pins the oracle's geometry conventions (view angle sign, detector orientation, fan
source position and magnification) against analytic line integrals of a Gaussian blob.

Blob f = exp(-|p-p0|^2/(2 sigma^2)).  Parallel ray (theta, t) with strip width w:
  strip integral = (pi sigma^2 / 1) [erf((t+w/2-c)/(sqrt2 sigma)) - erf((t-w/2-c)/(sqrt2 sigma))]
with c = x0 cos(theta) + y0 sin(theta).  Fan ray: the same with c replaced by the
signed distance of p0 from the ray line and w by the local strip width at p0."""
import numpy as np
import pytest
from scipy.special import erf


def blob(n, x0, y0, sig):
    c = np.arange(n) - n // 2
    X, Y = np.meshgrid(c, c)
    return np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2 * sig ** 2))


def strip_int(center, lo, hi, sig):
    # int_lo^hi sqrt(2 pi) sig exp(-(tau-center)^2/(2 sig^2)) dtau
    return np.pi * sig ** 2 * (erf((hi - center) / (np.sqrt(2) * sig)) -
                               erf((lo - center) / (np.sqrt(2) * sig)))


@pytest.mark.parametrize("d", [0.0, 300.0, 120.0])
def test_blob_projection(oracle_mod, d):
    n, W, npj, pf = 64, 95, 36, 72
    x0, y0, sig = 7.0, -4.0, 5.0
    img = blob(n, x0, y0, sig)
    cf = oracle_mod.conf(n, W, npj, pf, 1.0, 1.0, d, angle_mode=0)
    s, _ = oracle_mod.forward(cf, img)
    pred = np.zeros_like(s)
    for a in range(npj):
        th = a * 2 * np.pi / pf
        cT, sT = np.cos(th), np.sin(th)
        for b in range(W):
            t = (b - W // 2) * 1.0
            if d > 0:
                bq = np.hypot(d, t)
                cB, sB = d / bq, t / bq
                cR, sR = cB * cT - sB * sT, sB * cT + cB * sT
                dist = x0 * cR + y0 * sR - d * t / bq
                w = cB * (1 + (x0 * sT - y0 * cT) / d)
            else:
                dist = x0 * cT + y0 * sT - t
                w = 1.0
            pred[a, b] = strip_int(0.0, dist - w / 2, dist + w / 2, sig)
    # the blob is sampled at pixel centres: discretisation error ~ 1/(24 sigma^2)
    err = np.abs(s - pred).max() / pred.max()
    assert err < 5e-3, err
