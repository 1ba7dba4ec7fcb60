"""GPU parity of the projector pair (libctprj through ctypes) against the reference's
own cpuPrj (golden fixtures) and the FP64 oracle.

Bars (FP32 projection): forward and adjoint of smooth data <= 1e-5 relative L2 vs the
reference, with the error vs FP64 printed per case; adjoint of Gaussian noise <= 1e-5
relative to |A^T||y|; FBP <= 1e-4; <Ax,y> = <x,A^T y> <= 1e-6 relative; A and A^T
bitwise transposes of each other."""
import numpy as np
import pytest

from conftest import golden_cases, load_golden, rel_l2
from test_sym_decomposition import choose_group

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _proj(geo, flags=1):
    """flags: 0 general kernels, 1 symmetric where possible, 3 / 7 with the 8-fold
    (and 4-fold) group switched off."""
    from imgrecsrc_amd.projector import Projector
    P = Projector([geo], symmetric=bool(flags & 1))
    if flags not in (0, 1):
        P.set_tune(flags=flags)
    return P


def _expected_mode(geo, flags):
    return choose_group(geo[3], geo[2], geo[6] > 0, flags, n=geo[0])[0]


# FP64 oracle angle mode matching a slice's kernels: per-view float angles on the
# general kernels, the canonical view's angle mapped by the group on the symmetric ones
ANGLE_MODE = {0: 1, 1: 2, 2: 3, 3: 4}


def _run(P, x, op):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x, np.float32).ravel()).cuda()
    out = getattr(P, op)(t)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("case", golden_cases())
def test_golden_parity(torch_dev, oracle_mod, case):
    geo, z = load_golden(case)
    n, W, npj = geo[0], geo[1], geo[2]
    P = _proj(geo)
    cf = oracle_mod.conf(*geo)
    s = _run(P, z["fwd_in"], "forward").reshape(npj, W)
    s64, _ = oracle_mod.forward(cf, z["fwd_in"])
    e_ref, e_64 = rel_l2(s, z["fwd_out"]), rel_l2(s, s64)
    print(f"{case}: fwd vs reference {e_ref:.2e}, vs FP64 {e_64:.2e}")
    assert e_ref <= 1e-5 and e_64 <= 1e-5
    b = _run(P, z["bwd_in"], "adjoint").reshape(n, n)
    b64, _ = oracle_mod.adjoint(cf, z["bwd_in"])
    e_ref, e_64 = rel_l2(b, z["bwd_out"]), rel_l2(b, b64)
    print(f"{case}: adj vs reference {e_ref:.2e}, vs FP64 {e_64:.2e}")
    assert e_ref <= 1e-5 and e_64 <= 1e-5
    if "bwdg_in" in z:
        bg = _run(P, z["bwdg_in"], "adjoint").reshape(n, n)
        bg64, _ = oracle_mod.adjoint(cf, z["bwdg_in"])
        mag, _ = oracle_mod.adjoint(cf, np.abs(z["bwdg_in"]))
        e_ref = np.linalg.norm(bg - z["bwdg_out"]) / np.linalg.norm(mag)
        e_64 = np.linalg.norm(bg - bg64) / np.linalg.norm(mag)
        print(f"{case}: adj(noise) vs reference {e_ref:.2e}, vs FP64 {e_64:.2e} (of |A^T||y|)")
        assert e_ref <= 1e-5 and e_64 <= 2e-6
    if "fbp_in" in z:
        f = _run(P, z["fbp_in"], "fbp").reshape(n, n)
        f64 = oracle_mod.fbp(cf, z["fbp_in"])
        e_ref, e_64 = rel_l2(f, z["fbp_out"]), rel_l2(f, f64)
        print(f"{case}: FBP vs reference {e_ref:.2e}, vs FP64 {e_64:.2e}")
        assert e_ref <= 1e-4 and e_64 <= 1e-5
    # reference quirks
    if W % 2 == 0:
        assert np.all(s[:, 0] == 0)
    if n % 2 == 0:
        assert np.all(b[0, :] == 0) and np.all(b[:, 0] == 0)


TINY = [
    (8, 13, 12, 24, 1.0, 1.0, 0.0),
    (7, 10, 9, 18, 1.0, 1.0, 0.0),
    (6, 9, 10, 20, 1.3, 0.7, 0.0),
    (8, 12, 16, 16, 1.0, 1.0, 40.0),
    (9, 15, 11, 22, 0.8, 1.0, 30.0),
    (11, 16, 13, 13, 1.0, 0.9, 25.0),
    (12, 17, 24, 24, 1.0, 1.0, 30.0),      # fan, full turn, even n
    (9, 14, 20, 40, 1.0, 1.0, 0.0),        # parallel, half turn (folded)
    (10, 15, 28, 28, 1.0, 1.0, 0.0),       # parallel over a full turn
    (7, 12, 9, 24, 1.0, 1.0, 0.0),         # parallel, partial arc
    (9, 14, 18, 18, 1.0, 1.0, 30.0),       # fan, prjFull = 2 mod 4: 4-fold
    (8, 13, 9, 18, 1.0, 1.0, 0.0),         # parallel half turn, 4-fold folded
    (10, 15, 22, 22, 1.1, 0.9, 0.0),       # parallel full turn, 4-fold
    (7, 12, 13, 13, 1.0, 1.0, 25.0),       # fan, odd prjFull: 2-fold
    (7, 10, 11, 11, 1.0, 1.0, 0.0),        # parallel, odd prjFull: 2-fold
    (9, 13, 7, 13, 0.9, 1.1, 30.0),        # fan arc, odd prjFull: 2-fold, half used
    (8, 11, 6, 24, 1.0, 1.0, 20.0),        # fan quarter arc: 2-fold
]

# every geometry on the general and the automatically chosen kernels, and two with the
# larger groups switched off (4-fold and 2-fold on prjFull % 4 == 0, folded and not)
TRANSPOSE_CASES = ([(g, f) for g in TINY for f in (1, 0)] +
                   [((12, 17, 24, 24, 1.0, 1.0, 30.0), f) for f in (3, 7)] +
                   [((9, 14, 20, 40, 1.0, 1.0, 0.0), f) for f in (3, 7)])


@pytest.mark.parametrize("geo,flags", TRANSPOSE_CASES)
def test_forward_adjoint_bitwise_transposes(torch_dev, oracle_mod, geo, flags):
    """Columns of A from forward(e_p) and rows from adjoint(e_r): every weight is one
    term, so the two must agree bit for bit; and both within FP32 of the FP64 A at the
    angles the kernels use (per view on the general kernels, the canonical view's angle
    mapped by the symmetry group on the symmetric ones)."""
    import torch
    n, W, npj = geo[0], geo[1], geo[2]
    P = _proj(geo, flags)
    mode = _expected_mode(geo, flags)
    assert P.modes() == [mode]
    eye_img = torch.eye(n * n, device="cuda")
    eye_sino = torch.eye(npj * W, device="cuda")
    Af = torch.stack([P.forward(eye_img[p].contiguous()) for p in range(n * n)], 1)
    At = torch.stack([P.adjoint(eye_sino[r].contiguous()) for r in range(npj * W)], 0)
    torch.cuda.synchronize()
    Af, At = Af.cpu().numpy(), At.cpu().numpy()
    assert np.array_equal(Af, At)
    A64 = oracle_mod.dense(oracle_mod.conf(*geo, angle_mode=ANGLE_MODE[mode], support_mode=1))
    err = np.abs(Af - A64).max()
    print(f"{geo} flags={flags} mode={mode}: max |A - A64| = {err:.2e}")
    assert err <= 2e-6 * max(1.0, np.abs(A64).max())


@pytest.mark.parametrize("geo", [
    (256, 367, 180, 360, 1.0, 1.0, 0.0),          # 256^2 parallel workload
    (1024, 1024, 360, 360, 1.0, 1.0, 4096.0),     # 1024^2 fan workload
    (512, 729, 360, 360, 1.0, 1.0, 2000.0),       # 512^2 fan (beam hardening) workload
    (511, 730, 200, 720, 1.0, 1.0, 0.0),          # odd image, even detector, partial arc
])
def test_adjoint_identity(torch_dev, geo):
    import torch
    from imgrecsrc_amd.phantom import gaussian
    n, W, npj = geo[0], geo[1], geo[2]
    P = _proj(geo)
    x = torch.from_numpy(gaussian((n, n), 21)).cuda().reshape(-1)
    y = torch.from_numpy(gaussian((npj, W), 22)).cuda().reshape(-1)
    Ax = P.forward(x)
    Aty = P.adjoint(y)
    lhs = torch.dot(Ax.double(), y.double()).item()
    rhs = torch.dot(x.double(), Aty.double()).item()
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"{geo}: <Ax,y>={lhs:.9e} <x,A^T y>={rhs:.9e} rel={rel:.2e}")
    assert rel <= 1e-6


@pytest.mark.parametrize("geo,flags,mode", [
    ((256, 367, 180, 360, 1.0, 1.0, 0.0), 1, 1),
    ((257, 365, 360, 360, 1.0, 1.0, 1100.0), 1, 1),
    ((200, 290, 90, 360, 1.0, 1.0, 0.0), 1, 1),     # quarter turn: half the evaluations used
    ((256, 367, 181, 362, 1.0, 1.0, 0.0), 1, 2),    # parallel half turn, 4-fold folded
    ((300, 427, 298, 298, 1.0, 1.0, 1200.0), 1, 2),  # fan full turn, 4-fold
    ((255, 363, 361, 361, 1.0, 1.0, 1000.0), 1, 3),  # fan, odd prjFull: 2-fold
    ((257, 365, 360, 360, 1.0, 1.0, 1100.0), 3, 2),  # 8-fold geometry on the 4-fold group
    ((257, 365, 360, 360, 1.0, 1.0, 1100.0), 7, 3),  # ... and on the 2-fold group
])
def test_symmetric_matches_general(torch_dev, geo, flags, mode):
    """The symmetric kernels (weights at the canonical view) and the general kernels
    (weights at each view) agree to FP32 rounding of the view angles."""
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    n, W, npj = geo[:3]
    S, G = _proj(geo, flags), _proj(geo, 0)
    assert S.modes() == [mode] and G.modes() == [0]
    x = torch.from_numpy(ellipse_phantom(n, 9)).cuda().reshape(-1)
    y = S.forward(x)
    e_f = rel_l2(y.cpu().numpy(), G.forward(x).cpu().numpy())
    e_a = rel_l2(S.adjoint(y).cpu().numpy(), G.adjoint(y).cpu().numpy())
    print(f"{geo}: symmetric vs general fwd {e_f:.2e} adj {e_a:.2e}")
    assert e_f < 2e-6 and e_a < 2e-6


def test_batched_equals_single(torch_dev):
    """A ragged, mixed parallel/fan plan gives every slice bit-identical results to a
    one-slice plan of that slice."""
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    from imgrecsrc_amd.projector import Projector
    geoms = [(128, 185, 90, 180, 1.0, 1.0, 0.0), (96, 140, 101, 101, 1.0, 1.0, 384.0),
             (65, 96, 30, 120, 1.0, 1.0, 0.0), (200, 287, 77, 77, 1.0, 1.0, 800.0)]
    B = Projector(geoms)
    imgs = [torch.from_numpy(ellipse_phantom(g[0], 40 + i)).cuda() for i, g in enumerate(geoms)]
    sinos = [torch.from_numpy(gaussian((g[2], g[1]), 50 + i)).cuda() for i, g in enumerate(geoms)]
    sb = B.unpack_sinos(B.forward(B.pack_images(imgs)))
    ib = B.unpack_images(B.adjoint(B.pack_sinos(sinos)))
    for i, g in enumerate(geoms):
        S = Projector([g])
        assert torch.equal(sb[i].reshape(-1), S.forward(imgs[i].reshape(-1)))
        assert torch.equal(ib[i].reshape(-1), S.adjoint(sinos[i].reshape(-1)))


def test_work_order_does_not_change_results(torch_dev):
    """The work order (every slice spread over the XCDs, or small slices whole per XCD:
    tune flag bit 3) only schedules items; each item's sums are the same, so the outputs
    are bit-identical."""
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    from imgrecsrc_amd.projector import Projector
    geoms = [(96, 140, 101, 101, 1.0, 1.0, 384.0), (128, 185, 90, 180, 1.0, 1.0, 0.0),
             (65, 96, 30, 120, 1.0, 1.0, 0.0), (80, 117, 45, 90, 1.0, 1.0, 0.0),
             (70, 103, 70, 70, 1.0, 1.0, 280.0), (200, 287, 77, 77, 1.0, 1.0, 800.0)]
    B = Projector(geoms)
    imgs = [torch.from_numpy(ellipse_phantom(g[0], 60 + i)).cuda() for i, g in enumerate(geoms)]
    sinos = [torch.from_numpy(gaussian((g[2], g[1]), 70 + i)).cuda() for i, g in enumerate(geoms)]
    x, y = B.pack_images(imgs), B.pack_sinos(sinos)
    out = []
    for flags in (1, 9):
        B.set_tune(flags=flags)
        out.append((B.forward(x).clone(), B.adjoint(y).clone()))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


WINDOW_CASES = [(g, 1) for g in TINY] + [
    ((12, 17, 24, 24, 1.0, 1.0, 30.0), 3), ((12, 17, 24, 24, 1.0, 1.0, 30.0), 7),
    ((9, 14, 20, 40, 1.0, 1.0, 0.0), 3), ((9, 14, 20, 40, 1.0, 1.0, 0.0), 7),
    ((256, 367, 180, 360, 1.0, 1.0, 0.0), 1), ((257, 365, 360, 360, 1.0, 1.0, 1100.0), 1),
    ((300, 427, 298, 298, 1.0, 1.0, 1200.0), 1), ((255, 363, 361, 361, 1.0, 1.0, 1000.0), 1),
    ((64, 150, 60, 60, 0.45, 1.0, 200.0), 1)]  # fine pitch: several windows per pixel


@pytest.mark.parametrize("geo,flags", WINDOW_CASES)
def test_adjoint_window_width_is_bitwise_neutral(torch_dev, geo, flags):
    """The symmetric adjoint's bin windows (4 wide, or 3 wide with tune flag bit 4) only
    change how bins are loaded: every pixel still sums its bins in increasing order, so
    the adjoint and FBP are bit-identical."""
    import torch
    from imgrecsrc_amd.phantom import gaussian
    P = _proj(geo, flags)
    y = torch.from_numpy(gaussian((geo[2], geo[1]), 13)).cuda().reshape(-1)
    ref, fbp = P.adjoint(y).clone(), P.fbp(y).clone()
    P.set_tune(flags=flags | 16)
    assert torch.equal(P.adjoint(y), ref)
    assert torch.equal(P.fbp(y), fbp)


@pytest.mark.parametrize("geo,flags", WINDOW_CASES)
def test_adjoint_lds_staging_is_bitwise_neutral(torch_dev, geo, flags):
    """The symmetric adjoint's gathers through each wave's LDS area (tune flag bit 5:
    the wave stages the bins its pixels can read once per view and element) only change
    where the values come from: the same weights, the same order of sums, so the adjoint
    and FBP are bit-identical to the per-lane gathers, with 3- and 4-bin windows and
    with split work items; the fine-pitch case exceeds the staged width and falls back."""
    import torch
    from imgrecsrc_amd.phantom import gaussian
    P = _proj(geo, flags)
    y = torch.from_numpy(gaussian((geo[2], geo[1]), 17)).cuda().reshape(-1)
    for win, split in ((0, 1), (16, 1), (16, 3)):
        P.set_tune(flags=flags | win, adj_split=split)
        ref, fbp = P.adjoint(y).clone(), P.fbp(y).clone()
        P.set_tune(flags=flags | win | 32)
        assert torch.equal(P.adjoint(y), ref), (win, split)
        assert torch.equal(P.fbp(y), fbp), (win, split)
        P.set_tune(flags=flags | win, adj_split=1)


@pytest.mark.parametrize("geo,flags", WINDOW_CASES)
def test_split_work_items_match_one_work_group(torch_dev, geo, flags):
    """Work-groups per item (tune adj_split / fwd_split) only regroup the partial sums of
    the views (adjoint) or image rows (forward): forward, adjoint and FBP agree with the
    one-work-group results to FP32 rounding, and the forward keeps its exact zeros."""
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    P = _proj(geo, flags)
    x = torch.from_numpy(ellipse_phantom(geo[0], 5)).cuda().reshape(-1)
    y = torch.from_numpy(gaussian((geo[2], geo[1]), 14)).cuda().reshape(-1)
    f1, a1, b1 = P.forward(x).clone(), P.adjoint(y).clone(), P.fbp(y).clone()
    for k in (2, 3, 4):
        P.set_tune(adj_split=k, fwd_split=k)
        f, a, b = P.forward(x), P.adjoint(y), P.fbp(y)
        for got, ref in ((f, f1), (a, a1), (b, b1)):
            assert torch.allclose(got, ref, rtol=0, atol=2e-6 * ref.abs().max().item()), k
        assert torch.equal(f == 0, f1 == 0)
    P.set_tune(adj_split=1, fwd_split=1)
    assert torch.equal(P.adjoint(y), a1) and torch.equal(P.forward(x), f1)


def test_set_tune_changes_several_shapes_at_once(torch_dev):
    """One set_tune call that changes the general kernels' forward block and adjoint tile
    and the symmetric shapes together rebuilds every work list it affects."""
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian
    for geo, flags in (((64, 95, 36, 72, 1.0, 1.0, 0.0), 0), ((72, 103, 48, 48, 1.0, 1.0, 300.0), 1)):
        P = _proj(geo, flags)
        x = torch.from_numpy(ellipse_phantom(geo[0], 4)).cuda().reshape(-1)
        y = torch.from_numpy(gaussian((geo[2], geo[1]), 9)).cuda().reshape(-1)
        f1, a1 = P.forward(x).clone(), P.adjoint(y).clone()
        P.set_tune(fwd_block=64, adj_tile=8, adj_seg=2, fwd_seg=2, adj_split=2, fwd_split=3)
        assert torch.allclose(P.forward(x), f1, rtol=0, atol=2e-6 * f1.abs().max().item())
        assert torch.allclose(P.adjoint(y), a1, rtol=0, atol=2e-6 * a1.abs().max().item())


def test_split_work_items_in_a_mixed_batch(torch_dev):
    """Slices of the three symmetry groups in one plan: each group's launches use the same
    partial-sum space in stream order."""
    import torch
    from imgrecsrc_amd.phantom import gaussian
    from imgrecsrc_amd.projector import Projector
    geoms = [(64, 95, 36, 72, 1.0, 1.0, 0.0), (48, 70, 30, 30, 1.0, 1.0, 150.0),
             (40, 61, 25, 25, 1.0, 1.0, 0.0), (96, 139, 64, 64, 1.0, 1.0, 400.0)]
    P = Projector(geoms)
    assert sorted(set(P.modes())) == [1, 2, 3]
    x = torch.rand(P.img_elems, device="cuda")
    y = torch.from_numpy(gaussian((1, P.sino_elems), 15)).cuda().reshape(-1)
    f1, a1 = P.forward(x).clone(), P.adjoint(y).clone()
    P.set_tune(adj_split=3, fwd_split=2)
    assert torch.allclose(P.forward(x), f1, rtol=0, atol=2e-6 * f1.abs().max().item())
    assert torch.allclose(P.adjoint(y), a1, rtol=0, atol=2e-6 * a1.abs().max().item())


def test_nnz_matches_oracle(torch_dev, oracle_mod):
    """Device count of non-zero FP32 weights vs the FP64 count: overlaps of a pixel corner
    with a strip edge below ~1e-9 px^2 round to exactly 0 in FP32, so the FP32 count is
    slightly lower (by < 1e-3)."""
    for geo in [(64, 95, 36, 72, 1.0, 1.0, 0.0), (64, 95, 36, 72, 1.0, 1.0, 300.0)]:
        P = _proj(geo)
        g = P.nnz()
        o = oracle_mod.nnz(oracle_mod.conf(*geo, support_mode=1))
        assert o - 1e-3 * o <= g <= o, (g, o)


def test_mass_conservation_parallel(torch_dev):
    """Parallel beam, unit pitch: the strips of one view tile the plane, so every view
    of a phantom inside the detector's reach sums to the image mass."""
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom
    geo = (256, 367, 180, 360, 1.0, 1.0, 0.0)
    P = _proj(geo)
    img = ellipse_phantom(256, 1)
    s = _run(P, img, "forward").reshape(180, 367).astype(np.float64)
    # even n: row 0 and column 0 are outside the pixel grid the projector reads
    assert np.allclose(s.sum(1), img[1:, 1:].sum(dtype=np.float64), rtol=2e-6)


def test_plan_errors(torch_dev):
    from imgrecsrc_amd._lib import CtprjError
    from imgrecsrc_amd.projector import Projector
    with pytest.raises(CtprjError, match="inside the image"):
        Projector([(256, 367, 180, 360, 1.0, 1.0, 100.0)])
    P = _proj((16, 23, 8, 16, 1.0, 1.0, 0.0))
    import torch
    with pytest.raises(ValueError):
        P.forward(torch.zeros(15, device="cuda"))
    with pytest.raises(TypeError):
        P.forward(torch.zeros(256, device="cuda", dtype=torch.float64))
    before = P.get_tune()
    for bad in ({"adj_split": 9}, {"fwd_split": 0}, {"adj_seg": 16}, {"fwd_seg": 32},
                {"fwd_block": 96}, {"adj_tile": 32}, {"flags": 64}):
        with pytest.raises(CtprjError):
            P.set_tune(**bad)
    assert P.get_tune() == before  # a refused shape changes nothing


def test_autotune_roundtrip(torch_dev, tmp_path):
    import torch
    from imgrecsrc_amd.phantom import ellipse_phantom
    geo = (128, 185, 90, 180, 1.0, 1.0, 0.0)
    P = _proj(geo)
    x = torch.from_numpy(ellipse_phantom(128, 3)).cuda().reshape(-1)
    base = P.forward(x).clone()
    cache = str(tmp_path / "tune.txt")
    t = P.autotune(reps=2, cache_path=cache)
    assert t["fwd_block"] in (64, 128, 256) and t["adj_tile"] in (8, 16)
    assert t["adj_seg"] in (1, 2, 4, 8) and t["fwd_seg"] in (1, 2, 4, 8, 16)
    assert t["adj_split"] in (1, 2, 3, 4, 5, 6, 8) and t["fwd_split"] in (1, 2, 3, 4, 5, 6, 8)
    # shapes change the order of the partial sums, not the matrix
    assert torch.allclose(P.forward(x), base, rtol=1e-6, atol=1e-6 * base.abs().max().item())
    Q = _proj(geo)
    assert Q.autotune(reps=2, cache_path=cache) == t   # served from the cache file
