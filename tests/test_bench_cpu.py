"""CPU checks of bench.py's contract: the multi-GPU refusal and the roofline
arithmetic (no GPU is touched)."""
import importlib.util
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
