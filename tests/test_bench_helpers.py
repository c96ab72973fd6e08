"""The profile-summary helpers of bench.py (CPU only): HBM bytes per launch
from rocprofv3 FETCH_SIZE / WRITE_SIZE counter files, labelled kernel/limbs,
scaled by the device-copy calibration in the files or by given factors."""
import csv
import importlib.util
import os

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


FIELDS = ["Kernel_Name", "Grid_Size", "Workgroup_Size", "Counter_Name", "Counter_Value"]


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for name, grid, value in rows:
            w.writerow({"Kernel_Name": name, "Grid_Size": grid, "Workgroup_Size": 256,
                        "Counter_Name": counter, "Counter_Value": value})


GS = "void ck::(anonymous namespace)::gs_accum_kernel<8, 2>(ck::EwNp, ck::PrimeTab, ck::GsAccumArgs)"
BS = "ck::(anonymous namespace)::bs_fused_kernel(ck::EwNp, ck::PrimeTab, ck::BsFusedArgs)"
COPY = "__amd_rocclr_copyBuffer"


def test_pmc_traffic_labels_and_given_scale(tmp_path):
    b = _bench()
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    # gs_accum: 32768 threads per limb (35 limbs); bs_fused: 64 workgroups per
    # limb, limbs rounded up to a multiple of 8 (56)
    _write(f, "FETCH_SIZE", [(GS, 35 * 32768, 100.0), (GS, 35 * 32768, 300.0),
                             (BS, 56 * 64 * 256, 50.0)])
    _write(w, "WRITE_SIZE", [(GS, 35 * 32768, 10.0), (GS, 35 * 32768, 30.0),
                             (BS, 56 * 64 * 256, 5.0)])
    out = b.pmc_traffic(str(f), str(w), scale=(2.0, 1.0))
    k = out["kernels"]
    assert set(k) == {"gs_accum/35", "bs_fused/56"}
    assert k["gs_accum/35"]["fetch_kib"] == 200.0 and k["gs_accum/35"]["write_kib"] == 20.0
    assert k["gs_accum/35"]["hbm_bytes"] == int((200.0 * 2.0 + 20.0) * 1024)
    assert k["bs_fused/56"]["hbm_bytes"] == int((50.0 * 2.0 + 5.0) * 1024)


def test_pmc_traffic_calibration_copy(tmp_path):
    b = _bench()
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    # the 256 MiB copy (grid 131072) reads half its bytes in FETCH_SIZE units
    _write(f, "FETCH_SIZE", [(COPY, 131072, 131072.0), (GS, 8 * 32768, 64.0)])
    _write(w, "WRITE_SIZE", [(COPY, 131072, 262144.0), (GS, 8 * 32768, 16.0)])
    out = b.pmc_traffic(str(f), str(w))
    assert out["fetch_scale"] == 2.0 and out["write_scale"] == 1.0
    assert out["kernels"]["gs_accum/8"]["hbm_bytes"] == int((64.0 * 2.0 + 16.0) * 1024)
    # without the copy and without factors there is nothing to scale by
    _write(f, "FETCH_SIZE", [(GS, 8 * 32768, 64.0)])
    with pytest.raises(SystemExit):
        b.pmc_traffic(str(f), str(w))
