"""GPU parity of the wavelet transforms (rwt/mdwt_r.c, rwt/midwt_r.c) and the ADMM
wavelet prox (npg/admm.m) against the FP64 restatement.  The transforms run in FP64 in
the reference's summation order, so their FP32 outputs equal the rounded restatement bit
for bit; the ADMM solution agrees to 1e-6 of its scale (its stopping tests compare FP64
sums whose order differs)."""
import numpy as np
import pytest

from oracle import wavelet_oracle as W

pytestmark = pytest.mark.gpu

# levels whose block holds at most 4096 coefficients run in one LDS-resident work-group;
# 128^2, 256 x 64 and 256 x 128 start with one launch per level (one level of 128^2: no
# level fits in LDS); 8 x 4096 is wider than the one-launch level kernels and runs two
# passes per level
CASES = [(2, 3, 16, 16), (4, 3, 32, 16), (6, 4, 64, 64), (8, 2, 8, 12), (20, 2, 16, 16),
         (6, 6, 128, 128), (4, 5, 256, 64), (2, 7, 256, 128), (4, 2, 8, 4096), (4, 1, 128, 128)]


def _wv(N, L, R, C, batch=1):
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf
    h = daubcqf(N)[0]
    return Wavelet(R, C, h, L, batch=batch), h


@pytest.mark.parametrize("N,L,R,C", CASES)
def test_dwt_idwt_bitwise(N, L, R, C):
    import torch
    w, h = _wv(N, L, R, C)
    x = np.random.default_rng(N + L).standard_normal((R, C)).astype(np.float32)
    xt = torch.from_numpy(x).cuda().reshape(-1)
    y = w.mdwt(xt).cpu().numpy().reshape(R, C)
    yr = W.mdwt(x.astype(np.float64), h, L)
    np.testing.assert_array_equal(y, yr.astype(np.float32))
    z = w.midwt(torch.from_numpy(y).cuda().reshape(-1)).cpu().numpy().reshape(R, C)
    np.testing.assert_array_equal(z, W.midwt(y.astype(np.float64), h, L).astype(np.float32))
    np.testing.assert_allclose(z, x, atol=2e-6 * np.abs(x).max())
    # in place
    w.mdwt(xt, out=xt)
    np.testing.assert_array_equal(xt.cpu().numpy().reshape(R, C), y)


@pytest.mark.parametrize("n,L", [(32, 3), (128, 4)])
def test_batched_images_transform_independently(n, L):
    import torch
    w, h = _wv(4, L, n, n, batch=3)
    x = np.random.default_rng(9).standard_normal((3, n, n)).astype(np.float32)
    y = w.mdwt(torch.from_numpy(x).cuda().reshape(-1)).cpu().numpy().reshape(3, n, n)
    for b in range(3):
        np.testing.assert_array_equal(y[b], W.mdwt(x[b].astype(np.float64), h, L).astype(np.float32))
    z = w.midwt(torch.from_numpy(y).cuda().reshape(-1)).cpu().numpy().reshape(3, n, n)
    for b in range(3):
        np.testing.assert_array_equal(z[b], W.midwt(y[b].astype(np.float64), h, L).astype(np.float32))
    assert w.penalty(torch.from_numpy(x).cuda().reshape(-1)) == pytest.approx(
        sum(W.l1_penalty(x[b].astype(np.float64), h, L) for b in range(3)), rel=1e-12)


@pytest.mark.parametrize("N,L,shape,u,nonneg_active",
                         [(4, 3, (32, 32), 0.5, False), (6, 3, (64, 64), 2.0, True),
                          (2, 4, (32, 48), 1.0, True), (4, 4, (128, 128), 1.0, True),
                          (2, 2, (8, 4096), 1.0, True), (2, 1, (128, 128), 1.0, True)])
def test_admm_matches_oracle(N, L, shape, u, nonneg_active):
    import torch
    w, h = _wv(N, L, *shape)
    rng = np.random.default_rng(N * 7)
    a = rng.standard_normal(shape)
    if not nonneg_active:
        a = a + 10.0
    a = a.astype(np.float32)
    init = np.maximum(a, 0).astype(np.float32)
    out, itr = w.admm(torch.from_numpy(a).cuda().reshape(-1), u, relativeTol=1e-7, maxItr=3000,
                      init=torch.from_numpy(init).cuda().reshape(-1))
    ref, itr_ref = W.admm(a.astype(np.float64), u, h, L, relative_tol=1e-7, max_itr=3000,
                          init=init.astype(np.float64))
    got = out.cpu().numpy().reshape(shape)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"N={N} L={L} {shape}: iterations {itr} (oracle {itr_ref}), max rel diff {err:.2e}")
    assert err <= 1e-6
    assert abs(itr - itr_ref) <= max(2, itr_ref // 50)
    assert got.min() >= 0
    if not nonneg_active:
        closed = W.midwt(W.soft_thresh(W.mdwt(a.astype(np.float64), h, L), u), h, L)
        np.testing.assert_allclose(got, closed, atol=2e-5 * np.abs(a).max())


def test_admm_zero_input_and_bad_arguments():
    import torch
    w, h = _wv(2, 2, 8, 8)
    z = torch.zeros(64, dtype=torch.float32, device="cuda")
    out, itr = w.admm(z, 1.0)
    assert itr == 0 and not torch.any(out)
    # the loop state starts afresh on the next call
    a = np.random.default_rng(3).standard_normal((8, 8))
    out, itr = w.admm(torch.from_numpy(a.astype(np.float32)).cuda().reshape(-1), 0.3, 1e-7, 3000)
    ref, itr_ref = W.admm(a.astype(np.float32).astype(np.float64), 0.3, h, 2,
                          relative_tol=1e-7, max_itr=3000)
    assert np.abs(out.cpu().numpy().reshape(8, 8) - ref).max() <= 1e-6 * np.abs(ref).max()
    assert abs(itr - itr_ref) <= 2
    from imgrecsrc_amd._lib import CtprjError
    from imgrecsrc_amd.wavelet import Wavelet
    with pytest.raises(CtprjError):
        Wavelet(12, 12, [0.5, 0.5], 3)  # 12 not divisible by 8
    with pytest.raises(CtprjError):
        Wavelet(16, 16, [1.0, 0.5, 0.2], 2)  # odd filter length


def _disc(n, r):
    c = np.arange(n) - (n - 1) / 2
    X, Y = np.meshgrid(c, c)
    return X ** 2 + Y ** 2 <= r * r


def test_masked_admm_matches_oracle_vectors():
    """Psi / Psit with the supports opt.mask and maskk (tomography/ConfigCT.m:85-86): the
    device works on whole images with zeros outside, the restatement on the supports'
    vectors as the reference does."""
    import torch
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf, wvlt_mask
    h = daubcqf(6)[0]
    L, n = 3, 64
    m = _disc(n, 26)
    mk = wvlt_mask(m, h, L)
    w = Wavelet(n, n, h, L)
    w.set_masks(m, mk)
    a = np.random.default_rng(5).standard_normal((n, n)).astype(np.float32)
    out, itr = w.admm(torch.from_numpy(a).cuda().reshape(-1), 1.5, 1e-7, 3000)
    ops = W.masked_ops(h, L, m, mk)
    ref, itr_ref = W.admm(a.astype(np.float64)[m], 1.5, h, L, 1e-7, 3000, ops=ops)
    got = out.cpu().numpy().reshape(n, n)
    print(f"masked ADMM: iterations {itr} (oracle {itr_ref})")
    assert not got[~m].any()
    assert np.abs(got[m] - ref).max() <= 1e-6 * np.abs(ref).max()
    assert abs(itr - itr_ref) <= max(2, itr_ref // 50)
    x = torch.from_numpy(np.where(m, a, 0).astype(np.float32)).cuda().reshape(-1)
    assert w.penalty(x) == pytest.approx(np.abs(ops[1](a.astype(np.float64)[m])).sum(), rel=1e-9)
    # Psit returns the masked coefficients, Psi the masked image
    c = w.mdwt(x).cpu().numpy().reshape(n, n)
    assert not c[~mk].any()
    assert not w.midwt(torch.from_numpy(c).cuda().reshape(-1)).cpu().numpy().reshape(n, n)[~m].any()


def test_admm_on_one_context_across_mask_changes():
    """One context: an unmasked prox, then masks set, then other masks: every call
    matches its own restatement (the replayed iteration sequence follows the masks)."""
    import torch
    from imgrecsrc_amd.wavelet import Wavelet, daubcqf, wvlt_mask
    h = daubcqf(4)[0]
    L, n = 3, 128  # 128^2: one-launch levels and the LDS-resident coarse levels
    a = np.random.default_rng(8).standard_normal((n, n)).astype(np.float32)
    at = torch.from_numpy(a).cuda().reshape(-1)
    w = Wavelet(n, n, h, L)
    out, itr = w.admm(at, 1.2, 1e-7, 3000)
    ref, itr_ref = W.admm(a.astype(np.float64), 1.2, h, L, relative_tol=1e-7, max_itr=3000)
    assert np.abs(out.cpu().numpy().reshape(n, n) - ref).max() <= 1e-6 * np.abs(ref).max()
    assert abs(itr - itr_ref) <= max(2, itr_ref // 50)
    for r in (50, 37):
        m = _disc(n, r)
        mk = wvlt_mask(m, h, L)
        w.set_masks(m, mk)
        out, itr = w.admm(at, 1.2, 1e-7, 3000)
        ref, itr_ref = W.admm(a.astype(np.float64)[m], 1.2, h, L, 1e-7, 3000,
                              ops=W.masked_ops(h, L, m, mk))
        got = out.cpu().numpy().reshape(n, n)
        assert not got[~m].any()
        assert np.abs(got[m] - ref).max() <= 1e-6 * np.abs(ref).max(), r
        assert abs(itr - itr_ref) <= max(2, itr_ref // 50), r
