"""Seeded synthetic inputs for the CT operators (ellipse phantoms, Gaussian test vectors).

The ellipse phantom follows the workload definition used for every benchmark and parity
case: 10 ellipses with centres uniform in the disc of radius R = n/2 (r = R*sqrt(U),
phi = 2*pi*U), semi-axes U[0.05, 0.5]*R, rotation U[0, pi), additive amplitude U[0, 1],
rasterised at the projector's pixel centres x = col - n/2, y = row - n/2 (the image
layout of prj/cpuPrj.c:383).
"""
from __future__ import annotations

import numpy as np


def ellipse_params(seed: int, n: int, count: int = 10, amp_lo: float = 0.0,
                   amp_hi: float = 1.0) -> np.ndarray:
    """(count, 6) array of (x0, y0, a, b, rot, amp) in pixel units."""
    rng = np.random.Generator(np.random.PCG64(seed))
    R = n / 2.0
    out = np.zeros((count, 6))
    for i in range(count):
        r = R * np.sqrt(rng.uniform())
        ph = 2 * np.pi * rng.uniform()
        a = rng.uniform(0.05, 0.5) * R
        b = rng.uniform(0.05, 0.5) * R
        rot = rng.uniform(0, np.pi)
        amp = rng.uniform(amp_lo, amp_hi)
        out[i] = (r * np.cos(ph), r * np.sin(ph), a, b, rot, amp)
    return out


def rasterise(params: np.ndarray, n: int) -> np.ndarray:
    """float64 n x n image, row index = y + n/2, column index = x + n/2."""
    c = np.arange(n) - n // 2
    X, Y = np.meshgrid(c, c)  # X varies along columns, Y along rows
    img = np.zeros((n, n))
    for x0, y0, a, b, rot, amp in params:
        dx, dy = X - x0, Y - y0
        cr, sr = np.cos(rot), np.sin(rot)
        u = dx * cr + dy * sr
        v = -dx * sr + dy * cr
        img += amp * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return img


def ellipse_phantom(n: int, seed: int, dtype=np.float32) -> np.ndarray:
    return rasterise(ellipse_params(seed, n), n).astype(dtype)


def gaussian(shape, seed: int, dtype=np.float32) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(shape).astype(dtype)
