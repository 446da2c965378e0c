"""Host-side logic of bench.py (no GPU): the one-GPU contract and the HBM
traffic summary built from rocprofv3 counter files."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stdout


def _counter_csv(path, rows):
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(cols, r)))


def test_traffic_summary_scales_by_the_copy_probe(tmp_path, monkeypatch):
    out = tmp_path / "traffic.json"
    monkeypatch.setattr(bench, "PROFILE_TRAFFIC", str(out))
    gib = 2 ** 30
    # the copy probe: 1 GiB read and written; FETCH_SIZE reads half of it,
    # WRITE_SIZE all of it (values in KiB, split over two counter instances)
    _counter_csv(tmp_path / "cf.csv",
                 [(1, "fill_kernel", "FETCH_SIZE", 1.0),
                  (2, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib / 4096),
                  (2, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib / 4096)])
    _counter_csv(tmp_path / "cw.csv",
                 [(2, "__amd_rocclr_copyBuffer", "WRITE_SIZE", gib / 2048),
                  (2, "__amd_rocclr_copyBuffer", "WRITE_SIZE", gib / 2048)])
    # two trailing-update launches: 3 and 5 GiB fetched by the counter, 2 GiB
    # written; the panel-product instance and other kernels are not counted
    _counter_csv(tmp_path / "f.csv",
                 [(7, "update_kernel<false>(int)", "FETCH_SIZE", 3 * gib / 1024),
                  (9, "update_kernel<false>(int)", "FETCH_SIZE", 5 * gib / 1024),
                  (6, "update_kernel<true>(int)", "FETCH_SIZE", 1e9),
                  (8, "leaf_kernel", "FETCH_SIZE", 1e9)])
    _counter_csv(tmp_path / "w.csv",
                 [(7, "update_kernel<false>(int)", "WRITE_SIZE", 2 * gib / 1024),
                  (9, "update_kernel<false>(int)", "WRITE_SIZE", 2 * gib / 1024),
                  (6, "update_kernel<true>(int)", "WRITE_SIZE", 1e9)])
    bench.traffic_summary(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"),
                          str(tmp_path / "cf.csv"), str(tmp_path / "cw.csv"),
                          0, 896, gib, probe=(1024, 2048, 896))
    d = json.load(open(out))
    assert d["fetch_factor"] == pytest.approx(2.0)
    assert d["write_factor"] == pytest.approx(1.0)
    assert d["launches"] == 2
    assert d["fetch_bytes_per_launch"] == pytest.approx(8 * gib)   # (3+5)/2 x 2
    assert d["write_bytes_per_launch"] == pytest.approx(2 * gib)
    alg = 8.0 * (2 * 1024 * 2048 + 1024 * 896 + 2048 * 896)
    assert d["algorithmic_bytes_per_launch"] == pytest.approx(alg)
    assert d["update_probe"] == {"m": 1024, "n": 2048, "k": 896}
