"""The reference MEX call signature gpuPrj(arg0, conf, cmd) (prj/mPrj.c) on the GPU:
MATLAB column-major marshalling, double outputs, against the reference's golden
vectors."""
import numpy as np
import pytest

from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["par256", "fan128_even", "fan63"])
def test_gpuprj_signature(case):
    from imgrecsrc_amd import mex
    geo, z = load_golden(case)
    n, W, npj, pf, ds, er, d = geo
    conf = dict(n=n, prjWidth=W, np=npj, prjFull=pf, dSize=ds, effectiveRate=er, d=d)
    assert mex.gpuPrj(0, conf, "config") is None
    # MATLAB image matrix M(i,j) = C row i, column j
    s = mex.gpuPrj(z["fwd_in"].astype(np.float64), 0, "forward")
    assert s.shape == (W * npj, 1) and s.dtype == np.float64
    assert rel_l2(s.ravel(), z["fwd_out"].ravel()) < 1e-5
    # the same image as a column-major vector
    s2 = mex.gpuPrj(z["fwd_in"].reshape(-1, order="F"), 0, "forward")
    assert np.array_equal(s, s2)
    # backward takes the MATLAB prjWidth x np matrix (column a = view a)
    sino_matlab = z["bwd_in"].T.astype(np.float32)
    x = mex.gpuPrj(sino_matlab, 0, "backward")
    assert x.shape == (n * n, 1)
    img = x.reshape(n, n, order="F")
    assert rel_l2(img, z["bwd_out"]) < 1e-5
    if "fbp_in" in z:
        x = mex.gpuPrj(z["fbp_in"].T, 0, "FBP")
        assert rel_l2(x.reshape(n, n, order="F"), z["fbp_out"]) < 1e-4
    assert mex.gpuPrj(0, 0, "showConf") is None


def test_config_closures():
    """tomography/ConfigCT.m:109-135 closures with Ts scaling and a circular mask."""
    from imgrecsrc_amd import mex
    geo, z = load_golden("par65_even")
    n, W, npj, pf, ds, er, d = geo
    conf = dict(imgSize=n, prjWidth=W, prjNum=npj, prjFull=pf, dSize=ds, effectiveRate=er, dist=d)
    c = np.arange(n) - n // 2
    X, Y = np.meshgrid(c, c)
    mask = (X ** 2 + Y ** 2) <= (n // 2 - 1) ** 2
    Phi, Phit, FBP = mex.fan_par_ops(conf, Ts=0.5, mask=mask)
    xs = z["fwd_in"].astype(np.float64).reshape(-1, order="F")[np.flatnonzero(mask.reshape(-1, order="F"))]
    y = Phi(xs)
    full = np.where(mask, z["fwd_in"], 0)
    ref = mex.gpuPrj(full, 0, "forward") * 0.5
    assert np.allclose(y, ref)
    g = Phit(y)
    assert g.shape == (mask.sum(), 1)
    lhs = float(y.ravel() @ y.ravel())
    rhs = float(xs @ g.ravel())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
