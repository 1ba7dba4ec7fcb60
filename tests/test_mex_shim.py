"""The gpuPrj MEX gateway (imgrecsrc_amd/csrc/mex/gpuPrj_mex.c) built against a
stand-in MEX runtime (tests/mexmock): command set and marshalling of prj/mPrj.c:37-173,
MATLAB errors instead of silent returns.  CPU tests exercise the argument and state
errors; the GPU test runs 'config' / 'forward' / 'backward' / 'FBP' against the
reference's golden vectors."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, rel_l2

MOCK = os.path.join(ROOT, "tests", "mexmock")
SHIM = os.path.join(ROOT, "imgrecsrc_amd", "csrc", "mex", "gpuPrj_mex.c")


@pytest.fixture(scope="module")
def mex(tmp_path_factory):
    from imgrecsrc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from imgrecsrc_amd.build import build
        build()
    _lib.lib()  # torch's HIP runtime + libctprj first
    libdir = os.path.dirname(_lib.LIB_PATH)
    out = str(tmp_path_factory.mktemp("mex") / "libgpuprj_mexmock.so")
    subprocess.run(["gcc", "-std=c99", "-O1", "-shared", "-fPIC", "-Wall", "-Werror",
                    "-I", MOCK, "-I", os.path.join(ROOT, "include"), SHIM,
                    os.path.join(MOCK, "mexmock.c"), "-L", libdir, "-lctprj",
                    f"-Wl,-rpath,{libdir}", "-o", out], check=True)
    L = ctypes.CDLL(out)
    vp = ctypes.c_void_p
    L.mock_string.restype = vp
    L.mock_string.argtypes = [ctypes.c_char_p]
    L.mock_struct.restype = vp
    L.mock_set_field.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
    L.mxCreateNumericMatrix.restype = vp
    L.mxCreateNumericMatrix.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    L.mxGetPr.restype = ctypes.POINTER(ctypes.c_double)
    L.mxGetPr.argtypes = [vp]
    L.mxGetNumberOfElements.restype = ctypes.c_size_t
    L.mxGetNumberOfElements.argtypes = [vp]
    L.mock_call.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    L.mock_error_id.restype = ctypes.c_char_p
    L.mock_error_msg.restype = ctypes.c_char_p
    L.mock_free.argtypes = [vp]
    return L


def call(L, *args, nlhs=1):
    prhs = (ctypes.c_void_p * max(1, len(args)))(*args)
    plhs = (ctypes.c_void_p * 1)()
    rc = L.mock_call(nlhs, plhs, len(args), prhs)
    return rc, plhs[0], L.mock_error_id().decode(), L.mock_error_msg().decode()


def scalar(L, v):
    a = L.mxCreateNumericMatrix(1, 1, 6, 0)
    L.mxGetPr(a)[0] = v
    return a


def matrix(L, arr):
    """MATLAB double matrix with the given numpy 2-D array's element (i, j)."""
    arr = np.asarray(arr, np.float64)
    a = L.mxCreateNumericMatrix(arr.shape[0], arr.shape[1], 6, 0)
    flat = arr.reshape(-1, order="F")
    ctypes.memmove(L.mxGetPr(a), flat.ctypes.data, flat.nbytes)
    return a


def to_numpy(L, a):
    n = L.mxGetNumberOfElements(a)
    return np.ctypeslib.as_array(L.mxGetPr(a), shape=(n,)).copy()


def conf(L, geo):
    s = L.mock_struct()
    for k, v in zip(("n", "prjWidth", "np", "prjFull", "dSize", "effectiveRate", "d"), geo):
        L.mock_set_field(s, k.encode(), float(v))
    return s


def test_argument_errors(mex):
    rc, _, eid, _ = call(mex, scalar(mex, 0), scalar(mex, 0))
    assert rc == 1 and eid == "gpuPrj:nrhs"
    rc, _, eid, _ = call(mex, scalar(mex, 0), scalar(mex, 0), scalar(mex, 1))
    assert rc == 1 and eid == "gpuPrj:cmd"
    rc, _, eid, msg = call(mex, scalar(mex, 0), scalar(mex, 0), mex.mock_string(b"config"))
    assert rc == 1 and eid == "gpuPrj:config"
    s = mex.mock_struct()
    mex.mock_set_field(s, b"n", 8.0)
    rc, _, eid, msg = call(mex, scalar(mex, 0), s, mex.mock_string(b"config"))
    assert rc == 1 and eid == "gpuPrj:config" and "prjWidth" in msg


def test_device_error_is_a_matlab_error(mex):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    rc, _, eid, msg = call(mex, scalar(mex, 0), conf(mex, (16, 23, 8, 16, 1, 1, 0)),
                           mex.mock_string(b"config"))
    assert rc == 1 and eid == "gpuPrj:device" and "no HIP device" in msg


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["par256", "fan96_odd"])
def test_gateway_against_golden(mex, case):
    geo, z = load_golden(case)
    n, W, npj = geo[0], geo[1], geo[2]
    rc, _, eid, msg = call(mex, scalar(mex, 0), conf(mex, geo), mex.mock_string(b"config"))
    assert rc == 0, (eid, msg)
    rc, out, eid, msg = call(mex, matrix(mex, z["fwd_in"]), scalar(mex, 0), mex.mock_string(b"forward"))
    assert rc == 0, (eid, msg)
    s = to_numpy(mex, out)
    assert s.size == W * npj and rel_l2(s, z["fwd_out"].ravel()) < 1e-5
    rc, out, eid, msg = call(mex, matrix(mex, z["bwd_in"].T), scalar(mex, 0), mex.mock_string(b"backward"))
    assert rc == 0, (eid, msg)
    img = to_numpy(mex, out).reshape(n, n, order="F")
    assert rel_l2(img, z["bwd_out"]) < 1e-5
    rc, out, eid, msg = call(mex, matrix(mex, z["fbp_in"].T), scalar(mex, 0), mex.mock_string(b"FBP"))
    assert rc == 0, (eid, msg)
    assert rel_l2(to_numpy(mex, out).reshape(n, n, order="F"), z["fbp_out"]) < 1e-4
    rc, _, eid, _ = call(mex, matrix(mex, np.zeros((n, n + 1))), scalar(mex, 0), mex.mock_string(b"forward"))
    assert rc == 1 and eid == "gpuPrj:size"
