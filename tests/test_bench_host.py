"""bench.py's host-side pieces (no GPU): the HBM-traffic summary built from
rocprofv3 counter files, its reuse guard, and the one-GPU contract."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNEL = "_ZN3egs12_GLOBAL__N_118update192_w_kernelIDF16_Li2ELi2ELb0EEEvNS_7TileMapE"


def _counters(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)


def test_traffic_summary_calibrates_and_counts_the_schedule(tmp_path):
    d = tmp_path / "tr"
    # two dispatches of the kernel (one counter split over two rows, as
    # rocprofv3 writes per-instance values), one of another kernel, and the copy
    _counters(str(d / "f" / "f_counter_collection.csv"),
              [[1, KERNEL, "FETCH_SIZE", 1000.0], [1, KERNEL, "FETCH_SIZE", 1000.0], [2, KERNEL, "FETCH_SIZE", 4000.0],
               [3, "other_kernel", "FETCH_SIZE", 9e9]])
    _counters(str(d / "w" / "w_counter_collection.csv"), [[1, KERNEL, "WRITE_SIZE", 500.0], [2, KERNEL, "WRITE_SIZE", 1500.0]])
    _counters(str(d / "cf" / "cf_counter_collection.csv"), [[7, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 524288.0]])
    _counters(str(d / "cw" / "cw_counter_collection.csv"), [[7, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 1048576.0]])
    out = tmp_path / "t.json"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-summary", str(d),
                        "--traffic-out", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    t = json.loads(out.read_text())
    # a 1 GiB copy read as 512 MiB: fetches scale by 2, writes by 1
    assert abs(t["calibration"]["fetch"] - 2.0) < 1e-12 and abs(t["calibration"]["write"] - 1.0) < 1e-12
    assert t["dispatches"] == {"fetch": 2, "write": 2}
    assert abs(t["fetch_bytes_per_launch"] - 3000.0 * 1024 * 2) < 1e-6
    assert abs(t["write_bytes_per_launch"] - 1000.0 * 1024) < 1e-6
    # the default plan (4, 4) at n = 100,000, dts 960: 105 tile columns; the
    # profiler counted 882 dispatches of this kernel over three factorizations
    assert t["launches_per_factorization"] == 294 and 3 * 294 == 882
    assert t["tile_updates_per_factorization"] == 166650
    assert abs(t["algorithmic_bytes_per_launch"] - 166650 * 8.0 * 960 * 960 / 294) < 1.0


def test_committed_traffic_summary_matches_the_default_workload():
    sys.path.insert(0, ROOT)
    import argparse

    import bench

    a = argparse.Namespace(n=100000, dts=960, band_fp64=4, band_fp32=4, single_band=0, single_tau=2.0 ** -16,
                           fp64_product_band=None, plain_bf16=False, fp8_tau=None)
    traffic, alg = bench.measured_traffic("update_16_select", a)
    assert traffic is not None and 1.0 < traffic / alg < 3.0
    # another plan must not reuse it: other bands, or the FP8 tier off
    a.fp8_tau = 0.0
    assert bench.measured_traffic("update_16_select", a) == (None, None)
    a.fp8_tau = None
    a.band_fp64 = 2
    assert bench.measured_traffic("update_16_select", a) == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
