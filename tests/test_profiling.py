"""rocprofv3 CSV readers and the counter calibration behind the bench's ``roofline.traffic``
(CPU, synthetic CSV files in rocprofv3's column layout)."""
import csv

import pytest

from stenpy_amd import profiling


def _counters(path, rows):
    path.mkdir(parents=True, exist_ok=True)
    with open(path / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (kernel, counter, value) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": kernel, "Counter_Name": counter, "Counter_Value": value})


def test_traffic_scales_each_counter_by_the_copy_kernel(tmp_path):
    copy = 512 * 2 ** 20
    # FETCH_SIZE counts the copy's reads at half (KiB), WRITE_SIZE at face value
    _counters(tmp_path / "f", [("fused_kernel", "FETCH_SIZE", 3000.0), ("fused_kernel", "FETCH_SIZE", 1000.0),
                               ("fused_kernel", "FETCH_SIZE", 2000.0), ("other", "FETCH_SIZE", 9e9),
                               ("void k_unary_vec4<float>(int)", "FETCH_SIZE", copy / 1024 / 2)])
    _counters(tmp_path / "w", [("fused_kernel", "WRITE_SIZE", 500.0),
                               ("void k_unary_vec4<float>(int)", "WRITE_SIZE", copy / 1024)])
    t = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), kernel="fused_kernel",
                          copy_kernel="k_unary_", copy_bytes=copy)
    assert t["scale_fetch"] == pytest.approx(2.0) and t["scale_write"] == pytest.approx(1.0)
    # medians over dispatches: fetch 2000 KiB, write 500 KiB
    assert t["fetch_bytes"] == pytest.approx(2000 * 1024 * 2.0)
    assert t["bytes_per_launch"] == pytest.approx(2000 * 1024 * 2.0 + 500 * 1024)


def test_traffic_names_a_missing_kernel(tmp_path):
    _counters(tmp_path / "f", [("fused_kernel", "FETCH_SIZE", 1.0)])
    _counters(tmp_path / "w", [("fused_kernel", "WRITE_SIZE", 1.0)])
    with pytest.raises(KeyError, match="k_unary_"):
        profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), kernel="fused_kernel",
                          copy_kernel="k_unary_", copy_bytes=1)


def test_kernel_stats_and_markdown(tmp_path):
    with open(tmp_path / "run_kernel_stats.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"])
        w.writeheader()
        w.writerow({"Name": "fused_kernel", "Calls": 20, "TotalDurationNs": 14.5e6, "AverageNs": 725e3,
                    "Percentage": 90.0})
        w.writerow({"Name": "k_gen_fourier", "Calls": 2, "TotalDurationNs": 1.6e6, "AverageNs": 0.8e6,
                    "Percentage": 10.0})
    rows = profiling.kernel_stats(str(tmp_path))
    assert rows[0] == {"name": "fused_kernel", "calls": 20, "total_ns": 14.5e6, "avg_ns": 725e3, "pct": 90.0}
    md = profiling.stats_markdown(str(tmp_path), "cmd")
    assert md.splitlines()[4] == "| `fused_kernel` | 20 | 725.00 | 14.500 | 90.0 |"


def test_code_ops_per_voxel_counts_operators_and_helper_calls():
    from stenpy_amd.profiling import code_ops_per_voxel, _CALL_OPS
    lines = ["const float v4 = ((((qp0[0] + qm0[0]) - 2.0f * q00[0]) / 1.0f) + (((v0 + v1) - 2.0f * q00[0]) / 1.0f));",
             "const float v5 = tg_mid(a, b, 1.0f);", "float e47[3]; sym3_eigvalsh(t46, e47);",
             "const float v6 = -v5;"]                   # a unary minus is not an operation between operands
    assert code_ops_per_voxel(lines) == 9 + _CALL_OPS["tg_mid"] + _CALL_OPS["sym3_eigvalsh"]
