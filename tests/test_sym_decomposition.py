"""Index logic of the 8-fold symmetric kernels (imgrecsrc_amd/csrc/ctprj_sym.hip),
modelled on the CPU: every (view, bin, pixel) entry of A must be produced exactly once
by the forward decomposition and exactly once by the adjoint's orbit decomposition, and
both must name the same canonical weight W(a0, b, p) -- the condition for the two kernels
to be bitwise transposes.  Keys related by the exact half-turn identity
W(a0, b, p) == W(a0, 2pC-b, -p) of a folded parallel scan are identified; no other
geometric equivalence is assumed (at 45 degrees the diagonal mirror is not exact in FP32).
"""
from collections import defaultdict

import pytest


def apply_px(k, m, x, y):
    if m: y = -y
    k &= 3
    if k == 0: return x, y
    if k == 1: return -y, x
    if k == 2: return -x, -y
    return y, -x

def map_view(Q, pC, fold, k, m, a0):
    P = 4*Q
    v = ((2*Q - a0) if m else a0) + (k & 3)*Q
    v %= P
    sign = -1 if m else 1
    off = 2*pC if m else 0
    if fold and v >= 2*Q:
        v -= 2*Q; sign = -sign; off = 2*pC - off
    return v, sign, off

def canon(Q, np_, fold):
    out = []
    for a0 in range(Q//2 + 1):
        mirror = a0 == 0 or 2*a0 == Q
        ok = False
        for e in range(4 if mirror else 8):
            v, _, _ = map_view(Q, 0, fold, e & 3, e >> 2, a0)
            if v < np_: ok = True
        if ok: out.append(a0)
    return out

def decompose(n, W, np_, pf, fan):
    Q = pf // 4; pC = W//2; blo = pC - (W-1)//2; bhi = pC + (W-1)//2; H = (n-1)//2
    fold = (not fan) and np_ <= 2*Q
    cs = canon(Q, np_, fold)
    # canonical key normal form: (a0, b, p) ~ (a0, 2pC-b, -p) under fold
    def norm(a0, b, p):
        if fold:
            alt = (a0, 2*pC - b, (-p[0], -p[1]))
            return min((a0, b, p), alt)
        return (a0, b, p)
    fwd = {}
    bmax = pC if fold else bhi
    pixels = [(x, y) for y in range(-H, H+1) for x in range(-H, H+1)]
    for a0 in cs:
        mirror = a0 == 0 or 2*a0 == Q
        for b in range(blo, bmax+1):
            for p in pixels:
                for e in range(4 if mirror else 8):
                    v, sg, off = map_view(Q, pC, fold, e & 3, e >> 2, a0)
                    if v >= np_: continue
                    tgt = (v, off + sg*b, apply_px(e & 3, e >> 2, *p))
                    key = norm(a0, b, p)
                    if tgt in fwd and fwd[tgt] != key:
                        raise AssertionError(f"forward writes {tgt} twice: {fwd[tgt]} / {key}")
                    fwd[tgt] = key
    adj = defaultdict(list)
    for y0 in range(0, H+1):
        for x0 in range(y0, H+1):
            acc = [[] for _ in range(8)]
            for a0 in cs:
                mirror = a0 == 0 or 2*a0 == Q
                if mirror:
                    els = ([(0,0),(1,0),(0,1),(1,1)] if fold else [(k,m) for m in (0,1) for k in range(4)])
                else:
                    els = [(0,0),(1,0),(0,1),(3,1)] if fold else [(k,m) for m in (0,1) for k in range(4)]
                for (k2, m2) in els:
                    p = apply_px(k2, 1, x0, y0) if m2 else apply_px(4-k2, 0, x0, y0)
                    hs = ([h for h in range(8) if (h >> 2) == m2] if mirror else range(8))
                    for b in range(blo, bhi+1):
                        for h in hs:
                            k1, m1 = h & 3, h >> 2
                            k = k1 + (-k2 if m1 else k2); m = m1 ^ m2
                            v, sg, off = map_view(Q, pC, fold, k, m, a0)
                            if v >= np_: continue
                            acc[h].append(((v, off + sg*b), norm(a0, b, p)))
            nrot = 1 if (x0 == 0 and y0 == 0) else 4
            refl = 0 < y0 < x0
            for h in range(8):
                if (h < 4 and h >= nrot) or (h >= 4 and not refl): continue
                q = apply_px(h & 3, h >> 2, x0, y0)
                for (vb, key) in acc[h]:
                    adj[(vb[0], vb[1], q)].append(key)
    # compare
    bad = 0
    for tgt, keys in adj.items():
        if len(keys) != 1:
            bad += 1
        elif tgt not in fwd:
            bad += 1
        elif fwd[tgt] != keys[0]:
            bad += 1
    for tgt in fwd:
        if tgt not in adj:
            bad += 1
    # coverage: every (view<np, bin valid, pixel) present?
    tot = np_ * (bhi - blo + 1) * len(pixels)
    return len(fwd), len(adj), tot, bad



@pytest.mark.parametrize("geo", [
    (8, 13, 12, 24, False), (8, 12, 16, 16, True), (12, 17, 24, 24, True),
    (9, 14, 20, 40, False), (10, 15, 28, 28, False), (7, 12, 9, 24, False),
    (6, 9, 10, 20, False), (9, 12, 36, 36, True), (8, 11, 6, 24, True),
])
def test_each_entry_once_with_one_key(geo):
    nf, na, tot, bad = decompose(*geo)
    assert bad == 0
    assert nf == na == tot
