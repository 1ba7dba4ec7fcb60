"""The host C code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU):

* the gpuPrj MEX gateway (imgrecsrc_amd/csrc/mex/gpuPrj_mex.c) with the stand-in MEX
  runtime and a stand-in library, including a failed host staging allocation
  ('gpuPrj:nomem', forced by a malloc countdown);
* the FP64 oracle (oracle/ctprj_oracle.c) on small parallel and fan geometries.

Each driver is an executable built with gcc -fsanitize=address,undefined; a sanitizer
report makes it exit non-zero (halt_on_error)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "sanitize")
MOCK = os.path.join(ROOT, "tests", "mexmock")
INC = os.path.join(ROOT, "include")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")


def _build_run(tmp_path, name, cmds):
    for c in cmds:
        subprocess.run(c, check=True, cwd=tmp_path)
    r = subprocess.run([str(tmp_path / name)], capture_output=True, text=True, env=ENV,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Sanitizer" not in r.stderr, r.stderr
    return r.stdout


def test_gateway_under_sanitizers(tmp_path):
    shim = os.path.join(ROOT, "imgrecsrc_amd", "csrc", "mex", "gpuPrj_mex.c")
    out = _build_run(tmp_path, "shim_driver", [
        ["gcc", "-std=c99", *SAN, "-I", MOCK, "-I", INC, "-Dmalloc=mock_malloc", "-c", shim,
         "-o", "shim.o"],
        ["gcc", "-std=c99", *SAN, "-I", MOCK, "-I", INC, "-o", "shim_driver", "shim.o",
         os.path.join(HERE, "shim_driver.c"), os.path.join(HERE, "ctprj_stub.c"),
         os.path.join(HERE, "alloc_fail.c"), os.path.join(MOCK, "mexmock.c")],
    ])
    assert "all checks passed" in out


def test_oracle_under_sanitizers(tmp_path):
    src = os.path.join(ROOT, "oracle", "ctprj_oracle.c")
    out = _build_run(tmp_path, "oracle_driver", [
        ["gcc", "-std=c11", *SAN, "-pthread", "-I", os.path.join(ROOT, "oracle"), "-o",
         "oracle_driver", os.path.join(HERE, "oracle_driver.c"), src, "-lm"],
    ])
    assert "all checks passed" in out
