"""Pins the FP64 oracle against the reference's own cpuPrj (golden fixtures written by
tests/golden/make_golden.py from the reference's prj/cpuPrj.c, compiled by oracle/Makefile).

The reference is FP32; the oracle FP64.  Bars: forward / adjoint of smooth data 1e-6
relative L2; adjoint of Gaussian noise 5e-6 relative to the magnitude of the summed
terms (|A^T| |y|), since its own FP32 rounding grows with the number of terms; FBP
(high-pass ramp filter through a float FFT) 1e-4."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases, load_golden, rel_l2


def test_manifest_hashes():
    man = json.load(open(os.path.join(GOLDEN, "MANIFEST.json")))
    assert set(man["sha256"]) == {c + ".npz" for c in golden_cases()}
    for name, h in man["sha256"].items():
        with open(os.path.join(GOLDEN, name), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == h, name


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference(oracle_mod, case):
    geo, z = load_golden(case)
    cf = oracle_mod.conf(*geo)
    s, nnz = oracle_mod.forward(cf, z["fwd_in"])
    assert nnz > 0
    assert rel_l2(s, z["fwd_out"]) < 1e-6
    b, _ = oracle_mod.adjoint(cf, z["bwd_in"])
    assert rel_l2(b, z["bwd_out"]) < 1e-6
    if "bwdg_in" in z:
        bg, _ = oracle_mod.adjoint(cf, z["bwdg_in"])
        mag, _ = oracle_mod.adjoint(cf, np.abs(z["bwdg_in"]))
        err = np.linalg.norm(bg - z["bwdg_out"]) / np.linalg.norm(mag)
        assert err < 5e-6
    if "fbp_in" in z:
        fb = oracle_mod.fbp(cf, z["fbp_in"])
        assert rel_l2(fb, z["fbp_out"]) < 1e-4


@pytest.mark.parametrize("case", golden_cases())
def test_reference_quirks_in_fixtures(case):
    """Even detector: bin 0 is 0; even image: row 0 / column 0 of the adjoint are 0."""
    geo, z = load_golden(case)
    n, W = geo[0], geo[1]
    if W % 2 == 0:
        assert np.all(z["fwd_out"][:, 0] == 0)
    if n % 2 == 0:
        assert np.all(z["bwd_out"][0, :] == 0) and np.all(z["bwd_out"][:, 0] == 0)


def test_ramp_filter_restatement(oracle_mod):
    """The direct-convolution restatement of prj/utils.h:94-151 against an FFT
    evaluation of the same definition (zero-pad to 2N, ramp spectrum x Hann)."""
    rng = np.random.default_rng(3)
    for size in (1, 2, 7, 64, 129):
        sig = rng.standard_normal(size)
        N = 2 * size
        ramp = np.zeros(N)
        i = np.arange(1, N)
        kk = np.where(i < N // 2, i, N - i).astype(float)
        ramp[1:] = np.where(i % 2 == 1, -1.0 / (np.pi ** 2 * kk ** 2), 0.0)
        ramp[0] = 0.25
        R = np.fft.fft(ramp).real
        hann = 0.5 + 0.5 * np.cos(2 * np.pi * np.arange(N) / N)
        P = np.fft.fft(np.concatenate([sig, np.zeros(size)]))
        ref = np.fft.ifft(P * R * hann).real[:size]
        got = oracle_mod.ramp_filter(sig)
        assert np.allclose(got, ref, atol=1e-12, rtol=0)
