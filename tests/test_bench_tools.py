"""Host logic of bench.py's measurement tooling (no GPU): the summary that turns two
rocprofv3 --pmc runs into HBM bytes per operator call, and the one-GPU guard."""
import csv
import json
import os
import subprocess
import sys

import bench

FIELDS = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]
COPY_KIB = 256.0 * 1024.0  # the --calib-copy device copy


def _write(d, counter, rows):
    os.makedirs(d / "host", exist_ok=True)
    with open(d / "host" / "pmc_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (grid, name, val) in enumerate(rows):
            w.writerow({"Dispatch_Id": i + 1, "Grid_Size": grid, "Kernel_Name": name,
                        "Counter_Name": counter, "Counter_Value": val})


def test_pmc_summary_adds_the_kernels_of_one_call_at_the_timed_shape(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    (tmp_path / "profiles").mkdir()
    full = "void ctprj::k_adjoint_sym<false, 4, 1, 3>(ctprj::SliceDesc const*, float*)"
    # the autotuner's trial shapes (other grid sizes) come first, then the timed calls;
    # names arrive truncated (-T) or in full
    fetch = [(512, "k_adjoint_sym", 999.0), (1024, full, 888.0),
             (2048, "k_adjoint_sym", 100.0), (64, "k_adjoint_sym_combine", 40.0),
             (2048, full, 102.0), (64, "k_adjoint_sym_combine", 42.0),
             (256, "k_adjoint_sym_combine_other", 5000.0),
             (16777216, "__amd_rocclr_copyBuffer", 0.5 * COPY_KIB),
             (512, "__amd_rocclr_copyBuffer", 1.0)]
    write = [(2048, "k_adjoint_sym", 50.0), (64, "k_adjoint_sym_combine", 10.0),
             (16777216, "__amd_rocclr_copyBuffer", COPY_KIB)]
    _write(tmp_path / "f", "FETCH_SIZE", fetch)
    _write(tmp_path / "w", "WRITE_SIZE", write)
    bench.pmc_summary(str(tmp_path / "f"), str(tmp_path / "w"),
                      "k_adjoint_sym+k_adjoint_sym_combine", "adjoint_t", "cmd")
    out = json.load(open(tmp_path / "profiles" / "adjoint_t_pmc.json"))
    assert out["FETCH_SIZE_per_kernel"]["k_adjoint_sym"] == {"KiB": 101.0, "dispatches": 2,
                                                            "grid": 2048}
    assert out["FETCH_SIZE_per_kernel"]["k_adjoint_sym_combine"]["KiB"] == 41.0
    assert out["FETCH_SIZE_KiB"] == 142.0 and out["WRITE_SIZE_KiB"] == 60.0
    assert out["copy_fetch_factor"] == 0.5 and out["copy_write_factor"] == 1.0
    # bytes moved: each counter divided by what it reads per KiB of a plain copy
    assert out["traffic_bytes"] == (142.0 / 0.5 + 60.0) * 1024.0
    assert bench.traffic_from_profiles("adjoint_t") == out["traffic_bytes"]
    assert bench.traffic_from_profiles("absent") is None


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(bench.ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
