"""Pins the footprint weight: the oracle's restatement of prj/utils.h:12-92 against the
exact pixel/strip overlap area (polygon clipping), and the branch-free cumulative form
used by the HIP kernels against the same exact area."""
import numpy as np
import pytest


def _cases(rng, n):
    th = rng.uniform(0, 2 * np.pi, n)
    # exact axis / diagonal angles are special cases of the case split
    th[:8] = [0, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 4, 1e-9, np.pi / 4 + 1e-9, 5 * np.pi / 4]
    dist = rng.uniform(-2.5, 2.5, n)
    bw = rng.choice([0.25, 0.5, 0.8, 1.0, 1.5, 2.0, 3.0], n)
    return th, dist, bw


def test_get_weight_equals_strip_area(oracle_mod):
    rng = np.random.default_rng(11)
    th, dist, bw = _cases(rng, 4000)
    for t, d, w in zip(th, dist, bw):
        c, s = np.cos(t), np.sin(t)
        ref = oracle_mod.strip_area(d, w, c, s)
        got = oracle_mod.get_weight(d, w, c, s)
        assert abs(got - ref) <= 1e-12 * max(1.0, ref), (t, d, w, got, ref)


def test_strip_area_closed_values(oracle_mod):
    # theta = 0: strip of width 1 centred on the pixel covers it exactly
    assert oracle_mod.strip_area(0.0, 1.0, 1.0, 0.0) == pytest.approx(1.0, abs=1e-15)
    # half-overlap
    assert oracle_mod.strip_area(0.5, 1.0, 1.0, 0.0) == pytest.approx(0.5, abs=1e-15)
    # wide strip contains the pixel at any angle
    for t in np.linspace(0, np.pi, 13):
        assert oracle_mod.strip_area(0.1, 3.0, np.cos(t), np.sin(t)) == pytest.approx(1.0, abs=1e-13)
    # 45 degrees, centred, width w: the diagonal band minus the two corner triangles,
    # w*sqrt(2) - w^2/2
    w = 1e-3
    assert oracle_mod.strip_area(0.0, w, np.sqrt(0.5), np.sqrt(0.5)) == pytest.approx(
        w * np.sqrt(2) - w * w / 2, rel=1e-12)


def _gcum(u, b1, s):
    """The cumulative trapezoid integral of the HIP kernels (ctprj_device.h: gcum)."""
    b2 = b1 + s
    k = np.where(s > 0, 0.5 / np.where(s > 0, s, 1), 0.0)
    u1 = np.clip(u + b2, 0, s)
    u2 = np.clip(u + b1, 0, 2 * b1)
    u3 = np.clip(u - b1, 0, s)
    return (u2 + u3) + k * ((u1 - u3) * (u1 + u3))


def test_branch_free_form_equals_strip_area(oracle_mod):
    rng = np.random.default_rng(12)
    th, dist, bw = _cases(rng, 4000)
    for t, d, w in zip(th, dist, bw):
        c, s = abs(np.cos(t)), abs(np.sin(t))
        hi, lo = max(c, s), min(c, s)
        b1 = 0.5 * (hi - lo)
        got = (1 / hi) * (_gcum(w / 2 - d, b1, lo) - _gcum(-w / 2 - d, b1, lo))
        ref = oracle_mod.strip_area(d, w, np.cos(t), np.sin(t))
        assert abs(got - ref) <= 1e-12 * max(1.0, ref), (t, d, w, got, ref)


def test_weight_zero_outside_support(oracle_mod):
    # |dist| >= (|c|+|s|)/2 + bw/2 gives exactly zero: support bounds only need a superset
    rng = np.random.default_rng(13)
    for _ in range(500):
        t = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(t), np.sin(t)
        bw = rng.uniform(0.2, 2)
        edge = (abs(c) + abs(s)) / 2 + bw / 2
        d = edge * (1 + rng.uniform(0, 1)) * rng.choice([-1, 1])
        assert oracle_mod.get_weight(d, bw, c, s) == 0.0
