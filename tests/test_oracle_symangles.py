"""The FP64 oracle's symmetric angle modes (angle_mode 2/3/4): view a takes its canonical
view's float angle, rotated / reflected exactly by the group element that maps it.

Pins: (1) canonical views keep the reference's own angle, so their rows of A are
identical to angle_mode 1; (2) every other view lands on the geometrically right angle,
so A differs from angle_mode 1 only by the float rounding of the view angles (a wrong
rotation or reflection moves whole rows); (3) the reference's forward projector reuses
the canonical view's weights for the eight images (prj/cpuPrj.c:383-460), and the 8-fold
mode indeed reproduces its fan-beam outputs more closely than per-view angles do.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_l2

# (geometry, angle mode): prjFull % 4 == 0 for mode 2, even for mode 3, any for mode 4
CASES = [
    ((8, 13, 24, 24, 1.0, 1.0, 0.0), 2), ((9, 14, 24, 24, 1.0, 1.0, 30.0), 2),
    ((8, 13, 12, 24, 1.0, 1.0, 0.0), 2),
    ((9, 14, 18, 18, 1.0, 1.0, 30.0), 3), ((8, 13, 22, 22, 1.0, 1.0, 0.0), 3),
    ((9, 14, 24, 24, 1.0, 1.0, 30.0), 3),
    ((7, 12, 13, 13, 1.0, 1.0, 25.0), 4), ((9, 14, 24, 24, 1.0, 1.0, 30.0), 4),
    ((7, 10, 11, 11, 1.0, 1.0, 0.0), 4),
]


def canonical_views(P, mode):
    q = {2: 8, 3: 4, 4: 2}[mode]
    return [a for a in range(P) if q * a <= P]


@pytest.mark.parametrize("geo,mode", CASES)
def test_symmetric_angles(oracle_mod, geo, mode):
    n, W, npj, P = geo[:4]
    A1 = oracle_mod.dense(oracle_mod.conf(*geo, angle_mode=1, support_mode=1))
    As = oracle_mod.dense(oracle_mod.conf(*geo, angle_mode=mode, support_mode=1))
    A1 = A1.reshape(npj, W, n * n)
    As = As.reshape(npj, W, n * n)
    for a in canonical_views(P, mode):
        if a < npj:
            assert np.array_equal(As[a], A1[a]), a
    # other views: the same rays up to the float rounding of theta (|d theta| < 1e-6)
    diff = np.abs(As - A1).max()
    assert diff < 1e-5, diff
    # a wrong element is not a rounding-size change: mapping view a by theta -> -theta
    # in the 8-fold case moves rows by O(1)
    if mode == 2:
        wrong = As[:, ::-1, :]   # bins reversed: the reflection applied twice
        assert np.abs(wrong - A1).max() > 0.1


@pytest.mark.parametrize("case", ["fan128_even", "fan96_odd", "fan40_pitch", "fan4"])
def test_reference_forward_uses_canonical_angles(oracle_mod, case):
    geo, z = load_golden(case)
    e = {}
    for mode in (1, 2):
        s, _ = oracle_mod.forward(oracle_mod.conf(*geo, angle_mode=mode), z["fwd_in"])
        e[mode] = rel_l2(z["fwd_out"].ravel(), s.ravel())
    print(f"{case}: reference forward vs FP64 per-view angles {e[1]:.2e}, "
          f"canonical angles {e[2]:.2e}")
    assert e[2] < e[1]
