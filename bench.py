#!/usr/bin/env python3
"""Benchmark of the PBF substep on one MI355X.

Workload: the 1,048,576-particle dam break (128 x 64 x 128 lattice, spacing
0.5 h, h = 0.1, in a 12.8 x 6.4 x 6.4 box; rest density 1000, g = -9.8,
dt = 1/120 per substep, 4 Jacobi iterations, eps = 600, XSPH c = 0.01,
vorticity 0.0005, jitter +/- 0.01 h).  A "step" is one substep: the whole
body of Simulation/Source/main.cpp:71-124 through ``pbf_step``.

Prints one JSON line.  ``value`` is particle-substeps/s with the particles
already resident in HBM.  ``roofline`` reports the dominant kernel's
algorithmic bytes per launch over its average launch time (HIP events on the
library's stream, taken inside the timed region).  ``cpu_baseline`` times the
CPU oracle (a restatement of the reference's OpenMP path) on a bounded sample
of the same scene.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-substeps/s @1M, <=1% density err: plain vs learned warm-start"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle for each kernel launch: the compulsory
# per-particle streams of that pass (positions and velocities as 12-byte
# Vec<float,3>, scalars 4 bytes); neighbour gathers are assumed served on chip
# and neighbour lists are not counted.
ALG_BYTES = {
    "density_lambda": 12 + 4 + 4,          # R x*; W rho, lambda
    "dp_update": 12 + 4 + 12,              # R x*, lambda; W x*
    "fill_neighbors": 12 + 4,              # R x*; W count (lists not counted)
    "viscosity": 12 + 12 + 4 + 12,         # R x*, v, rho; W v
    "viscosity_positions": 12 + 12 + 4 + 12 + 12,
    "vorticity": 12 + 12 + 12,             # R x*, v; W omega
    "apply_vorticity": 12 + 12 + 12 + 12,  # R x*, omega, v; W v
    "forces_predict": 12 + 12 + 12 + 12,   # R x, v; W v, x*
    "bin": 12 + 4 + 4,
    "sort": 4 * 4,
    "reorder": 2 * (3 * 12 + 4 + 4 + 12 + 4 + 4),
    "scan": 8,
    "lambda_repair": 0,
}


def _load(name):
    return importlib.import_module(name)


def cpu_baseline(params, pos, vel, warm=1, timed=5):
    """Oracle in the reference's in-place mode with OpenMP on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    O = _load("oracle_lib")
    threads = max(1, min(16, os.cpu_count() or 1))
    o = O.Oracle(O.params_from_pbf(params), mode=O.SERIAL, threads=threads)
    o.set_state(pos, pos, vel)
    o.step(warm)
    times = []
    for _ in range(timed):
        t0 = time.perf_counter()
        o.step(1)
        times.append(time.perf_counter() - t0)
    o.close()
    med = statistics.median(times)
    n = pos.shape[0]
    return {
        "value": n / med,
        "unit": "particle-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} particles (same scene), {warm} warm-up + {timed} timed substeps, median "
                  f"{med:.3f} s/substep; oracle in the reference's OpenMP in-place mode",
    }


def _profile_traffic(kernel: str):
    """HBM bytes per launch from committed rocprofv3 PMC summaries, if present."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


KERNELS = ["forces_predict", "bin", "scan", "sort", "reorder", "fill_neighbors", "density_lambda", "mark_dirty",
           "lambda_repair", "dp_update", "viscosity", "vorticity", "apply_vorticity", "viscosity_positions"]


def timed_run(pbf, params, pos, vel, steps, profile=True):
    """K substeps from the initial scene on a fresh context; HIP events around
    every kernel launch on the library's stream when profile is set."""
    n = pos.shape[0]
    g = pbf.Particles(params)
    g.add(pos, pos, vel)
    if profile:
        g.profile_kernels(KERNELS)
    g.sync()
    t0 = time.perf_counter()
    g.region_begin()
    g.step(steps)
    ev_ms = g.region_end()
    g.sync()
    wall = time.perf_counter() - t0
    kern = {}
    if profile:
        for k in KERNELS:
            launches, ms = g.kernel_time(k)
            kern[k] = {"launches": launches, "total_ms": ms}
    st = g.stats()
    res = {"value": n * steps / wall, "seconds": wall, "event_ms": ev_ms, "kernels": kern, "cells": g.num_cells(),
           "stats": {"neighbor_overflow": st["neighbor_overflow"], "density_hacks": st["density_hacks"],
                     "max_density_error": round(st["max_density_error"], 5)}}
    g.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--lattice", default="128x64x128", help="block size in particles (x by y by z)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--kernel", default="", help="kernel for the roofline (default: the slowest in the timed run)")
    ap.add_argument("--cap", type=int, default=60, help="neighbour list cap (MAX_NEIGHBORS = 60)")
    ap.add_argument("--no-variants", dest="variants", action="store_false", help="skip the cap-128 variant")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB mebibytes device to device 5 times (PMC byte-counter calibration)")
    args = ap.parse_args(argv)

    if args.calibrate_copy:
        pbf = _load("sph-cpp_amd")
        nbytes = args.calibrate_copy << 20
        a, b = pbf.DeviceBuffer(nbytes), pbf.DeviceBuffer(nbytes)
        for _ in range(5):
            pbf._check(pbf.lib().pbf_memcpy_device(b.ptr, a.ptr, nbytes))
        print(json.dumps({"copy_bytes": nbytes, "copies": 5}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the single-domain PBF step has no multi-GPU path in this build "
              "(see the sharding design notes).", file=sys.stderr)
        return 2

    pbf = _load("sph-cpp_amd")
    scenes = _load("sph-cpp_amd.scenes")
    counts = tuple(int(c) for c in args.lattice.lower().split("x"))
    params, pos, vel = scenes.config_c2(seed=args.seed, lattice_counts=counts, cap=args.cap)
    n = pos.shape[0]

    # warm-up on a throwaway context (module load, first launches)
    w = pbf.Particles(params)
    w.add(pos, pos, vel)
    w.step(args.warmup)
    w.sync()
    w.close()

    res = timed_run(pbf, params, pos, vel, args.steps)
    dom = args.kernel or max(res["kernels"], key=lambda k: res["kernels"][k]["total_ms"])
    value = res["value"]
    kd = res["kernels"][dom]
    avg_launch_ms = kd["total_ms"] / max(1, kd["launches"])
    alg = ALG_BYTES.get(dom, 0) * n
    achieved = alg / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    roof = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": _profile_traffic(dom),
        "alg_bytes_per_launch": alg,
        "avg_launch_ms": round(avg_launch_ms, 5),
        "launches": kd["launches"],
    }
    # whole-substep bookkeeping: B_alg = 260 + 44 I + 8 C/N bytes per particle-substep
    I = params.solve_step_count
    b_alg = 260 + 44 * I + 8.0 * res["cells"] / n
    step_roof = {"alg_bytes_per_particle_substep": round(b_alg, 2),
                 "achieved_GBs": round(b_alg * value / 1e9, 2),
                 "frac_of_hbm_peak": round(b_alg * value / 1e9 / HBM_PEAK_GBS, 5)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain PBF dam break, {n} particles ({args.lattice} lattice, spacing 0.05, h 0.1) in "
                        f"12.8x6.4x6.4, dt 1/120, {I} Jacobi iterations, eps 600, XSPH 0.01, vorticity 0.0005, "
                        f"substeps 1..{args.steps} from the initial scene",
            "particles": n,
            "solver": "plain, fixed iterations",
            "jacobi_iterations": I,
            "neighbor_cap": int(params.max_neighbors),
            "cells": res["cells"],
        },
        "roofline": roof,
        "step_roofline": step_roof,
        "kernels_ms_per_step": {k: round(v["total_ms"] / args.steps, 4) for k, v in res["kernels"].items()},
        "stats_last_step": res["stats"],
    }
    if args.variants:
        vparams, _, _ = scenes.config_c2(seed=args.seed, lattice_counts=counts, cap=128)
        vr = timed_run(pbf, vparams, pos, vel, args.steps, profile=False)
        out["variants"] = {"cap128": {"value": vr["value"], "ms_per_step": vr["seconds"] * 1e3 / args.steps,
                                      "stats_last_step": vr["stats"],
                                      "note": "same scene with a 128-entry neighbour list (no list saturation)"}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(params, pos, vel)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
