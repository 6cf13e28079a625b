"""CPU checks of bench.py's contract: the multi-GPU refusal, the roofline
arithmetic and the profile summaries made from rocprofv3's CSV files (no GPU
is touched)."""
import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_kernel_roofline_arithmetic(bench):
    """achieved = algorithmic bytes per launch / mean launch time; frac = achieved / 8 TB/s."""
    n = 1 << 20
    prof = {"kernels": {"density_lambda": {"launches": 800, "total_ms": 100.0},
                        "fill_heavy": {"launches": 200, "total_ms": 65.0},
                        "scan": {"launches": 0, "total_ms": 0.0}}}
    kr = bench.kernels_roofline(prof, n)
    assert set(kr) == {"density_lambda", "fill_heavy"}
    d = kr["density_lambda"]
    assert d["alg_bytes_per_launch"] == 20 * n
    assert d["avg_launch_ms"] == pytest.approx(0.125)
    assert d["achieved_GBs"] == pytest.approx(20 * n / 0.125e-3 / 1e9, abs=0.1)
    assert d["hbm_frac"] == pytest.approx(d["achieved_GBs"] / 8000.0, abs=1e-4)
    assert kr["fill_heavy"]["achieved_GBs"] == 0.0


def test_step_algorithmic_figures(bench):
    """B_alg = 260 + 44 I + 8 C/N bytes and F_alg = 9 K + 324 + 56 (90 I + 125) operations per
    particle-substep: 438.6 B and 30,841 at the 1M dam break (I = 4, C/N = 0.326, K = 373)."""
    I, cells, n = 4, 341824, 1 << 20
    assert 260 + 44 * I + 8.0 * cells / n == pytest.approx(438.6, abs=0.05)
    assert 9 * bench.F_ALG_K + 324 + bench.F_ALG_NBR * (90 * I + 125) == 30841
    assert bench.VALU_LANE_OPS_PEAK == pytest.approx(3.93216e13)


def _csv(path, header, rows):
    import csv
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)  # kernel names hold commas: quoted as rocprofv3 quotes them
        w.writerow(header)
        w.writerows(rows)


def test_profile_summaries_from_rocprof_csv(bench, tmp_path):
    """summarize() on hand-made rocprofv3 CSVs: only the launches after the
    last particle upload (the timed run) count, the HBM counters are scaled by
    a 512 MiB copy's reading, and VALU issue = instructions / (SIMDs x mean
    duration x clock)."""
    d = str(tmp_path / "prof")
    k_den, k_up = "void pbf::k_density_lambda_c<0, false, true>(unsigned int)", "void pbf::k_widen3(float*)"
    # trace: a density launch before the upload (not timed), the upload, then two timed launches of 90 and 110 us
    trace_rows = [(k_den, 0, 500_000), (k_up, 600_000, 601_000), (k_den, 700_000, 790_000),
                  (k_den, 800_000, 910_000)]
    _csv(os.path.join(d, "trace", "c2_kernel_trace.csv"), ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
         trace_rows)
    _csv(os.path.join(d, "trace", "c2_kernel_stats.csv"),
         ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs"],
         [(k_den, 3, 700000, 233333, 99.9, 90000, 500000)])

    def counters(path, dispatches):
        # one row per counter of each dispatch, all with the dispatch's id
        rows = [(did, kern, name, val) for did, (kern, vals) in enumerate(dispatches) for name, val in vals.items()]
        _csv(path, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], rows)

    # FETCH_SIZE reads 2x the true KiB (the copy of 512 MiB reads 1 GiB), WRITE_SIZE 1x
    counters(os.path.join(d, "pmc_fetch", "c2_counter_collection.csv"),
             [(k_den, {"FETCH_SIZE": 1.0}), (k_up, {"FETCH_SIZE": 1.0}), (k_den, {"FETCH_SIZE": 200.0}),
              (k_den, {"FETCH_SIZE": 400.0})])
    counters(os.path.join(d, "pmc_write", "c2_counter_collection.csv"),
             [(k_den, {"WRITE_SIZE": 1.0}), (k_up, {"WRITE_SIZE": 1.0}), (k_den, {"WRITE_SIZE": 100.0}),
              (k_den, {"WRITE_SIZE": 100.0})])
    counters(os.path.join(d, "cal_fetch", "cp_counter_collection.csv"),
             [("copy", {"FETCH_SIZE": 2 * 512 * 1024.0})])
    counters(os.path.join(d, "cal_write", "cp_counter_collection.csv"),
             [("copy", {"WRITE_SIZE": 512 * 1024.0})])

    def sq(valu):
        return {"SQ_WAVES": 100.0, "SQ_INSTS_VALU": valu, "SQ_INSTS_VMEM_RD": 5000.0, "SQ_INSTS_LDS": 0.0,
                "SQ_ACTIVE_INST_VALU": 1000.0, "SQ_THREAD_CYCLES_VALU": 48000.0}
    counters(os.path.join(d, "pmc_sq", "c2_counter_collection.csv"),
             [(k_den, sq(1.0)), (k_up, sq(1.0)), (k_den, sq(216000.0)), (k_den, sq(216000.0))])

    out = str(tmp_path / "out")
    bench.summarize(d, "w", out)
    kern_md = open(os.path.join(out, "w_kernels.md")).read()
    assert "| `pbf::k_density_lambda_c<0, false, true>` | 2 | 100.0 | 90.0 | 110.0 | 0.20 | 100.0 |" in kern_md
    traffic = json.load(open(os.path.join(out, "pmc_traffic.json")))
    assert traffic["fetch_scale"] == pytest.approx(0.5) and traffic["write_scale"] == pytest.approx(1.0)
    t = traffic["pbf::k_density_lambda_c<0, false, true>"]
    # mean FETCH 300 KiB x 0.5, WRITE 100 KiB x 1
    assert t["read_bytes"] == 150 * 1024 and t["write_bytes"] == 100 * 1024 and t["launches"] == 2
    valu = json.load(open(os.path.join(out, "pmc_valu.json")))
    v = valu["pbf::k_density_lambda_c<0, false, true>"]
    # 216,000 instructions / (1024 SIMDs x 100 us x 2.4 GHz) = 0.000879; 2,160 per wave; lanes 48000 / (64 x 1000)
    assert v["valu_instr_per_wave"] == 2160
    assert v["valu_issue_per_simd_cycle"] == pytest.approx(216000 / (1024 * 100e-6 * 2.4e9), abs=1e-4)
    assert v["vector_loads_per_wave"] == 50.0 and v["valu_lanes_busy"] == pytest.approx(0.75)
