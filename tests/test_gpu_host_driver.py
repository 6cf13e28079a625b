"""The C++ host driver (sph-cpp_amd/host/sph_main.cpp) restating Simulation/Source/main.cpp's
step loop over the C ABI: per-pass calls, the fused step and the graph-replayed step write
identical position records, in the AdiosWriter record layout (adios_writer.h:78-114)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "sph-cpp_amd", "sph_run")


def read_records(path):
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        glob, local, offset = np.frombuffer(data, dtype=np.int64, count=3, offset=off)
        off += 24
        recs.append((int(glob), int(local), int(offset),
                     np.frombuffer(data, dtype=np.float32, count=local // 4, offset=off).reshape(-1, 3)))
        off += int(local)
    return recs


def test_host_driver_per_pass_equals_fused(tmp_path):
    assert os.path.exists(RUNNER), "build with python sph-cpp_amd/build.py"
    outs = {}
    for mode in ("per-pass", "fused", "graph"):
        path = str(tmp_path / f"{mode}.raw")
        args = [RUNNER, "--lattice", "16x12x10", "--box", "1.6", "0.8", "0.8", "--steps", "6", "--every", "2",
                "--write", path] + {"per-pass": [], "fused": ["--fused"], "graph": ["--graph"]}[mode]
        res = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        print(res.stdout.strip(), flush=True)
        outs[mode] = read_records(path)
    a, b, g = outs["per-pass"], outs["fused"], outs["graph"]
    assert len(a) == len(b) == len(g) == 3
    n = 16 * 12 * 10
    for ra, rb, rg in zip(a, b, g):
        assert ra[:3] == (n * 12, n * 12, 0)
        assert np.array_equal(ra[3], rb[3])
        assert np.array_equal(ra[3], rg[3])
        assert np.all(np.isfinite(ra[3]))


def test_host_driver_tolerance_solve(tmp_path):
    res = subprocess.run([RUNNER, "--lattice", "16x12x10", "--box", "1.6", "0.8", "0.8", "--steps", "6",
                          "--tolerance", "0.05"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    print(res.stdout.strip(), flush=True)
    assert '"mode": "solve"' in res.stdout
