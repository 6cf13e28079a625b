"""GPU tests of PBF_OPT_FAST_MATH: the fused step's neighbour sums with the
hardware square root and reciprocal instead of the correctly rounded ones.

Cell keys, particle order and neighbour lists do not depend on that
arithmetic (the list radius is tested on squared distances), so they stay
bit-exact.  Positions and velocities are compared with the exact mode, which
equals the oracle bit for bit (tests/test_gpu_parity.py), and the size of
the difference is set against the spread between the oracle in float and
the same step rerun in double (the reference at Real = double):

* after one substep of the 32,768-particle dam break the largest position
  difference to the exact mode is at most FAST_VS_EXACT_1 (stated below, in
  units of h) and at most 2x the float-vs-double spread;
* after 100 substeps (the run has reached the neighbour-cap collapse, where
  trajectories are chaotic) the RMS position difference to the exact mode is
  at most 2x the RMS float-vs-double spread, and the centre of mass and the
  kinetic energy agree with the exact mode to within 2x their own
  float-vs-double differences (plus a floor of 1e-6 relative).
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
H = 0.1
FAST_VS_EXACT_1 = 1e-5  # max |dx| / h after one substep


def _params(pbf, params, fast):
    p = pbf.Params.from_buffer_copy(params)
    if fast:
        p.options |= pbf.PBF_OPT_FAST_MATH
    return p


def _run(pbf, params, pos, vel, steps, fast):
    g = pbf.Particles(_params(pbf, params, fast))
    g.add(pos, pos, vel)
    g.step(steps)
    out = {"x": g.positions(), "v": g.velocities(), "lists": g.neighbor_lists(), "keys": g.bin_ids(),
           "ids": g.particle_ids()}
    g.close()
    return out


def _oracle64(params, pos, vel, steps):
    o = O.Oracle(O.params_from_pbf(params), double=True, mode=O.JACOBI, threads=THREADS)
    o.set_state(pos.astype(np.float64), pos.astype(np.float64), vel.astype(np.float64))
    o.step(steps)
    st = o.state()
    o.close()
    return st


def _rms(a, b):
    return float(np.sqrt(np.mean(np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2, axis=1))))


def test_fast_math_one_substep(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=3)
    e = _run(pbf, params, pos, vel, 1, False)
    f = _run(pbf, params, pos, vel, 1, True)
    for k in ("keys", "ids"):
        assert np.array_equal(e[k], f[k]), k
    for a, b in zip(e["lists"], f["lists"]):
        assert np.array_equal(a, b), "neighbour lists"
    d = _oracle64(params, pos, vel, 1)
    fe = float(np.abs(f["x"].astype(np.float64) - e["x"]).max()) / H
    ed = float(np.abs(e["x"].astype(np.float64) - d["x"]).max()) / H
    ve = float(np.abs(f["v"].astype(np.float64) - e["v"]).max())
    vd = float(np.abs(e["v"].astype(np.float64) - d["v"]).max())
    print(f"fast math, 1 substep: max |dx|/h fast-exact {fe:.3e}, exact-fp64 {ed:.3e}; "
          f"max |dv| fast-exact {ve:.3e}, exact-fp64 {vd:.3e}", flush=True)
    assert fe <= FAST_VS_EXACT_1
    assert fe <= 2.0 * ed
    assert ve <= 2.0 * vd


def test_fast_math_100_substeps_statistics(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=3)
    e = _run(pbf, params, pos, vel, 100, False)
    f = _run(pbf, params, pos, vel, 100, True)
    print("fast math: oracle fp64 rerun of 100 substeps", flush=True)
    d = _oracle64(params, pos, vel, 100)
    fe, ed = _rms(f["x"], e["x"]) / H, _rms(e["x"], d["x"]) / H
    com = [np.mean(a.astype(np.float64), axis=0) for a in (e["x"], f["x"], d["x"])]
    ke = [0.5 * float(np.sum(a.astype(np.float64) ** 2)) for a in (e["v"], f["v"], d["v"])]
    com_fe, com_ed = float(np.abs(com[1] - com[0]).max()), float(np.abs(com[0] - com[2]).max())
    ke_fe, ke_ed = abs(ke[1] - ke[0]) / ke[0], abs(ke[0] - ke[2]) / ke[0]
    print(f"fast math, 100 substeps: RMS |dx|/h fast-exact {fe:.3e}, exact-fp64 {ed:.3e}; "
          f"centre of mass {com_fe:.3e} vs {com_ed:.3e}; kinetic energy {ke_fe:.3e} vs {ke_ed:.3e}", flush=True)
    assert fe <= 2.0 * ed
    assert com_fe <= 2.0 * com_ed + 1e-6 * float(np.abs(com[0]).max())
    assert ke_fe <= 2.0 * ke_ed + 1e-6


def test_fast_math_solve_graph_runs_to_tolerance(pbf, scenes):
    """The iterate-to-tolerance solve in fast mode meets the same criterion;
    its slot counts stay close to the exact mode's."""
    params, pos, vel = scenes.config_c1(seed=3)
    params.options |= pbf.PBF_OPT_SOLVE_MEAN_ERROR
    slots = {}
    for fast in (False, True):
        g = pbf.Particles(_params(pbf, params, fast))
        g.add(pos, pos, vel)
        g.solve_graph(0.01, 64, pbf.PBF_WARM_PREVIOUS, 20)
        log = g.solve_log(20)
        g.close()
        assert len(log) == 20
        assert all(r["mean_error"] <= np.float32(0.01) or r["iterations"] == 64 for r in log)
        slots[fast] = np.array([r["iterations"] for r in log])
    print(f"fast math solve: slots exact {slots[False].sum()}, fast {slots[True].sum()}", flush=True)
    assert abs(int(slots[True].sum()) - int(slots[False].sum())) <= 0.1 * slots[False].sum() + 20
