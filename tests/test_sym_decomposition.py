"""Index logic of the symmetric kernels (imgrecsrc_amd/csrc/ctprj_sym.hip, group choice
in ctprj_api.hip), modelled on the CPU for the three symmetry groups D4 / V4 / Z2.

Every (view, bin, pixel) entry of A must be produced exactly once by the forward
decomposition and exactly once by the adjoint's orbit decomposition, and both must name
the same canonical weight W(a0, b, p) -- the condition for the two kernels to be bitwise
transposes.  Keys related by the exact half-turn identity W(a0, b, p) == W(a0, 2pC-b, -p)
of a folded parallel scan are identified; no other geometric equivalence is assumed (at
45 degrees the diagonal mirror is not exact in FP32).  A planted mistake in the model
(all elements on a mirror view, a missing fold element, a hole in the fundamental
domain, a dropped dedup, two writers of a folded scan's centre bin or of an even
detector's empty bin 0) makes these checks
fail, which the planted-mistake test
confirms.  The model checks the decomposition's consistency only: that each view map
is the right geometric one is pinned by the GPU comparison against the FP64 oracle.
"""
from collections import defaultdict

import pytest

D4, V4, Z2 = 1, 2, 3
# (all, rot, fold, mfold_rot, mfold_ref, q) -- group_masks() of ctprj_device.h
MASKS = {D4: (0xFF, 0x0F, 0x93, 0x03, 0x30, 8),
         V4: (0x55, 0x05, 0x11, 0x01, 0x10, 4),
         Z2: (0x41, 0x01, 0x00, 0x00, 0x00, 2)}


def bits(mask):
    return [e for e in range(8) if mask >> e & 1]


def apply_px(k, m, x, y):
    if m: y = -y
    k &= 3
    if k == 0: return x, y
    if k == 1: return -y, x
    if k == 2: return -x, -y
    return y, -x


def map_view(P, pC, fold, k, m, a0):
    v4 = (-4 * a0 if m else 4 * a0) + ((k + 2 * m) & 3) * P
    assert v4 % 4 == 0, "element outside the slice's group"
    v = (v4 >> 2) % P
    sign = -1 if m else 1
    off = 2 * pC if m else 0
    if fold and 2 * v >= P:
        v -= P >> 1; sign = -sign; off = 2 * pC - off
    return v, sign, off


def symmetric_views(P, np_, grp, fold):
    allm, rot, _, _, _, q = MASKS[grp]
    canon, slots = [], 0
    a0 = 0
    while q * a0 <= P:
        mirror = a0 == 0 or q * a0 == P
        els = bits(rot if mirror else allm)
        vs = [map_view(P, 0, fold, e & 3, e >> 2, a0)[0] for e in els]
        if min(vs) < np_:
            canon.append(a0)
            slots += len(els) // 2 if fold else len(els)
        a0 += 1
    return canon, (np_ / slots if slots else 0.0)


def choose_group(P, np_, fan, flags=1, n=None):
    """The host's choice (ctprj_api.hip: choose_group); images under 5x5 (n given) stay
    on the general kernels."""
    if not flags & 1 or np_ > P or (n is not None and n < 5):
        return 0, False, []
    for grp in (D4, V4, Z2):
        if grp == D4 and (P % 4 or flags & 2): continue
        if grp == V4 and (P % 2 or flags & 4): continue
        fold = (not fan) and grp != Z2 and 2 * np_ <= P
        canon, use = symmetric_views(P, np_, grp, fold)
        if use >= 0.5:
            return grp, fold, canon
    return 0, False, []


def decompose(n, W, np_, P, fan, flags=1, plant=None):
    grp, fold, cs = choose_group(P, np_, fan, flags)
    assert grp, "geometry not symmetric"
    allm, rot, gfold, mrot, mref, q = MASKS[grp]
    ref = allm & ~rot
    pC = W // 2; blo = pC - (W - 1) // 2; bhi = pC + (W - 1) // 2; H = (n - 1) // 2
    mv = map_view

    def norm(a0, b, p):
        if fold:
            return min((a0, b, p), (a0, 2 * pC - b, (-p[0], -p[1])))
        return (a0, b, p)

    def mirror(a0):
        return a0 == 0 or q * a0 == P

    # forward: one canonical ray (a0, b) and pixel p, images g p for g in the mask
    fwd = {}
    bmax = pC if fold else bhi
    pixels = [(x, y) for y in range(-H, H + 1) for x in range(-H, H + 1)]
    writes = defaultdict(int)
    for a0 in cs:
        els = bits(rot if mirror(a0) and plant != "mirror" else allm)
        # items start at bin 0: b = 0 < blo on an even detector stores the empty bin 0
        for b in range(0, bmax + 1):
            for e in els:
                v, sg, off = mv(P, pC, fold, e & 3, e >> 2, a0)
                if v >= np_: continue
                # the kernel's single writer of a folded scan's centre bin and bin 0
                if fold and (e & 2) and plant != "centre" and (b == pC or (b < blo and plant != "zero")):
                    continue
                if b < blo:
                    writes[(v, 0)] += 1
                    continue
                writes[(v, off + sg * b)] += 1
                for p in pixels:
                    tgt = (v, off + sg * b, apply_px(e & 3, e >> 2, *p))
                    key = norm(a0, b, p)
                    if tgt in fwd and fwd[tgt] != key:
                        raise AssertionError(f"forward writes {tgt} twice: {fwd[tgt]} / {key}")
                    fwd[tgt] = key
    # every sinogram entry is written by exactly one (work item, image)
    if any(c != 1 for c in writes.values()):
        raise AssertionError("a sinogram entry has more than one writer")
    # adjoint: orbit base pixels in the group's fundamental domain (8x8 tiles, as the
    # host lays them out), element sets per view kind, first-occurrence writes
    nt = (H + 1 + 7) // 8
    if grp == D4:
        tiles = [(ty * 8, tx * 8) for ty in range(nt) for tx in range(ty, nt)]
    elif grp == V4:
        tiles = [(ty * 8, tx * 8) for ty in range(nt) for tx in range(nt)]
    else:
        tiles = [(y, tx * 8) for y in range(-H, H + 1, 8) for tx in range(nt)]
    adj = defaultdict(list)
    for (r0, c0) in tiles:
        for lane in range(64):
            x0, y0 = c0 + (lane & 7), r0 + (lane >> 3)
            if not (x0 <= H and y0 <= (x0 if grp == D4 else H)):
                continue
            if plant == "domain" and x0 == 0:
                continue
            acc = [[] for _ in range(8)]
            for a0 in cs:
                if mirror(a0):
                    sets = [(mrot if fold else rot, rot), (mref if fold else ref, ref)]
                else:
                    sets = [(gfold if fold else allm, allm)]
                if plant == "fold" and fold and not mirror(a0):
                    sets = [(gfold & ~(1 << max(bits(gfold))), allm)]
                for gset, hmask in sets:
                    for e2 in bits(gset):
                        k2, m2 = e2 & 3, e2 >> 2
                        p = apply_px(k2, 1, x0, y0) if m2 else apply_px(4 - k2, 0, x0, y0)
                        for b in range(blo, bhi + 1):
                            for h in bits(hmask):
                                k1, m1 = h & 3, h >> 2
                                k = k1 + (-k2 if m1 else k2); m = m1 ^ m2
                                v, sg, off = mv(P, pC, fold, k, m, a0)
                                if v >= np_: continue
                                acc[h].append(((v, off + sg * b), norm(a0, b, p)))
            seen = []
            for h in bits(allm):
                qh = apply_px(h & 3, h >> 2, x0, y0)
                if qh in seen and plant != "dedup":
                    continue
                seen.append(qh)
                for (vb, key) in acc[h]:
                    adj[(vb[0], vb[1], qh)].append(key)
    bad = sum(1 for tgt, keys in adj.items()
              if len(keys) != 1 or tgt not in fwd or fwd[tgt] != keys[0])
    bad += sum(1 for tgt in fwd if tgt not in adj)
    tot = np_ * (bhi - blo + 1) * len(pixels)
    return grp, fold, len(fwd), len(adj), tot, bad


# (n, W, np, prjFull, fan, flags, expected group, expected fold)
CASES = [
    (8, 13, 12, 24, False, 1, D4, True), (8, 12, 16, 16, True, 1, D4, False),
    (12, 17, 24, 24, True, 1, D4, False), (9, 14, 20, 40, False, 1, D4, True),
    (10, 15, 28, 28, False, 1, D4, False), (7, 12, 9, 24, False, 1, D4, True),
    (6, 9, 10, 20, False, 1, D4, True), (9, 12, 36, 36, True, 1, D4, False),
    (8, 11, 6, 24, True, 1, Z2, False),   # a 90-degree arc: D4 only 1/4 useful
    # prjFull = 2 mod 4: the 4-fold group
    (9, 14, 18, 18, True, 1, V4, False), (9, 14, 10, 18, True, 1, V4, False),
    (8, 13, 9, 18, False, 1, V4, True), (8, 13, 14, 18, False, 1, V4, False),
    (10, 15, 22, 22, False, 1, V4, False), (8, 13, 11, 22, False, 1, V4, True),
    (6, 9, 5, 10, False, 1, V4, True),
    # odd prjFull: the 2-fold group
    (7, 12, 13, 13, True, 1, Z2, False), (9, 12, 7, 13, True, 1, Z2, False),
    (7, 10, 11, 11, False, 1, Z2, False), (10, 13, 15, 15, True, 1, Z2, False),
    # prjFull % 4 == 0 with D4 (and V4) switched off by the tune flags
    (8, 13, 16, 16, True, 3, V4, False), (9, 14, 12, 24, False, 3, V4, True),
    (8, 12, 16, 16, True, 7, Z2, False), (9, 13, 12, 24, False, 7, Z2, False),
]


@pytest.mark.parametrize("case", CASES)
def test_each_entry_once_with_one_key(case):
    n, W, np_, P, fan, flags, grp_exp, fold_exp = case
    grp, fold, nf, na, tot, bad = decompose(n, W, np_, P, fan, flags)
    assert (grp, fold) == (grp_exp, fold_exp)
    assert bad == 0
    assert nf == na == tot


@pytest.mark.parametrize("plant,case", [
    ("centre", (9, 14, 20, 40, False, 1)), ("centre", (8, 13, 9, 18, False, 1)),
    ("mirror", (8, 12, 16, 16, True, 1)), ("mirror", (7, 12, 13, 13, True, 1)),
    ("domain", (9, 14, 18, 18, True, 1)), ("fold", (9, 14, 20, 40, False, 1)),
    ("fold", (8, 13, 9, 18, False, 1)), ("dedup", (7, 12, 13, 13, True, 1)),
    ("dedup", (9, 14, 18, 18, True, 1)), ("zero", (9, 14, 20, 40, False, 1)),
])
def test_planted_mistakes_are_caught(plant, case):
    try:
        _, _, nf, na, tot, bad = decompose(*case, plant=plant)
    except AssertionError:
        return
    assert bad > 0 or not (nf == na == tot)


def test_group_choice():
    # full fan turns: the largest integral group
    assert choose_group(360, 360, True)[0] == D4
    assert choose_group(362, 362, True)[0] == V4
    assert choose_group(361, 361, True)[0] == Z2
    # parallel half turns fold (D4/V4); an odd prjFull cannot fold and Z2 is half useful
    assert choose_group(360, 180, False)[:2] == (D4, True)
    assert choose_group(362, 181, False)[:2] == (V4, True)
    assert choose_group(361, 180, False)[:2] == (Z2, False)
    # a quarter-turn arc: D4/V4 images land 1/4 on computed views, Z2's just over 1/2
    assert choose_group(360, 90, True)[0] == Z2
    assert symmetric_views(360, 90, D4, False)[1] == 0.25
    # more views than prjFull, or the symmetric kernels switched off
    assert choose_group(360, 361, True)[0] == 0
    assert choose_group(360, 360, True, flags=0)[0] == 0
    assert choose_group(8, 8, True, n=4)[0] == 0 and choose_group(8, 8, True, n=5)[0] == D4
