"""GPU tests of pbf_autotune_layout, the tuner of the bin spacing (cell size
and list radius) and the list cap.

* The tuner leaves the particles as it found them and applies the layout it
  reports: one step of the tuned context equals, bit for bit, one step of a
  fresh context created with the tuned parameters from the same state (single
  and batched contexts; in a batched context every scene's grid is derived
  again from its box).
* A tuned layout lists every neighbour within h (no list reaches its cap), so
  the sums run over the same particles as the reference's spacing with an
  uncapped list: after one substep of the 32,768-particle dam break the
  positions differ from that run only by summation order, at most
  LAYOUT_VS_UNCAPPED_1 (in units of h) and at most 2x the run's own
  float-vs-double spread (the oracle rerun in double).
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
H = 0.1
LAYOUT_VS_UNCAPPED_1 = 1e-5  # max |dx| / h after one substep


def _state(pbf, params, pos, vel, steps):
    g = pbf.Particles(params)
    g.add(pos, pos, vel)
    g.step(steps)
    x, v = g.positions(), g.velocities()
    g.close()
    return x, v


def test_layout_tuner_applies_its_layout_and_restores_the_state(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=4, cap=128)
    x, v = _state(pbf, params, pos, vel, 5)
    g = pbf.Particles(params)
    g.add(x, x, v)
    lay = g.autotune_layout(2)
    print(f"layout: spacing {lay['neighbor_bin_spacing']:.4f}, cap {lay['max_neighbors']}, "
          f"longest list {lay['longest_list']}, {lay['ms_per_step']:.3f} ms/step", flush=True)
    tuned = g.params
    assert np.float32(tuned.neighbor_bin_spacing) == np.float32(lay["neighbor_bin_spacing"])
    assert tuned.max_neighbors == lay["max_neighbors"]
    assert lay["longest_list"] <= lay["max_neighbors"] <= 128
    assert tuned.neighbor_bin_spacing >= np.float32(H)
    assert np.array_equal(g.positions(), x) and np.array_equal(g.velocities(), v), "state not restored"
    g.step(1)
    fresh = pbf.Params.from_buffer_copy(tuned)
    f = pbf.Particles(fresh)
    f.add(x, x, v)
    f.step(1)
    assert np.array_equal(g.positions(), f.positions())
    assert np.array_equal(g.velocities(), f.velocities())
    g.close()
    f.close()


def test_layout_tuner_sums_over_the_same_neighbours(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=4, cap=128)
    g = pbf.Particles(params)
    g.add(pos, pos, vel)
    g.autotune_layout(1)
    tuned = pbf.Params.from_buffer_copy(g.params)
    tuned.options |= pbf.PBF_OPT_COUNT_OVERFLOW
    g.set_params(tuned)
    g.step(1)
    assert g.stats()["neighbor_overflow"] == 0
    xt = g.positions()
    g.close()
    u = pbf.Particles(params)  # the reference's spacing, cap 128: no list is cut at this state
    uc = pbf.Params.from_buffer_copy(params)
    uc.options |= pbf.PBF_OPT_COUNT_OVERFLOW
    u.set_params(uc)
    u.add(pos, pos, vel)
    u.step(1)
    assert u.stats()["neighbor_overflow"] == 0
    xu = u.positions()
    u.close()
    o = O.Oracle(O.params_from_pbf(params), double=True, mode=O.JACOBI, threads=THREADS)
    o.set_state(pos.astype(np.float64), pos.astype(np.float64), vel.astype(np.float64))
    o.step(1)
    xd = o.state()["x"]
    o.close()
    tu = float(np.abs(xt.astype(np.float64) - xu).max()) / H
    ud = float(np.abs(xu.astype(np.float64) - xd).max()) / H
    print(f"layout {tuned.neighbor_bin_spacing:.4f}/{tuned.max_neighbors}: max |dx|/h tuned-uncapped {tu:.3e}, "
          f"uncapped-fp64 {ud:.3e}", flush=True)
    assert tu <= LAYOUT_VS_UNCAPPED_1
    assert tu <= 2.0 * ud


def test_layout_tuner_batched_scenes(pbf, scenes):
    params, boxes, sc = scenes.batched_dam_breaks(num_scenes=6, seed=3, nmin=2048, nmax=8192)
    params.max_neighbors = 128
    g = pbf.Particles(params, boxes=boxes)
    for k, (p, v) in enumerate(sc):
        g.add_to_scene(k, p, p, v)
    g.step(3)
    lay = g.autotune_layout(1)
    print(f"batched layout: spacing {lay['neighbor_bin_spacing']:.4f}, cap {lay['max_neighbors']}", flush=True)
    x, v = g.positions(), g.velocities()
    g.step(2)
    f = pbf.Particles(pbf.Params.from_buffer_copy(g.params), boxes=boxes)
    # the same particles, scene by scene, in the same order
    start = 0
    for k, (p, _) in enumerate(sc):
        m = p.shape[0]
        f.add_to_scene(k, x[start:start + m], x[start:start + m], v[start:start + m])
        start += m
    f.step(2)
    assert np.array_equal(g.positions(), f.positions())
    assert np.array_equal(g.velocities(), f.velocities())
    g.close()
    f.close()


def test_set_params_cap_limit(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=4, cap=64)
    g = pbf.Particles(params)
    g.add(pos, pos, vel)
    bigger = pbf.Params.from_buffer_copy(params)
    bigger.max_neighbors = 72
    with pytest.raises(pbf.PBFError):
        g.set_params(bigger)
    smaller = pbf.Params.from_buffer_copy(params)
    smaller.max_neighbors = 48
    g.set_params(smaller)
    g.step(1)
    lists, counts = g.neighbor_lists()
    assert lists.shape[1] == 48 and int(counts.max()) <= 48
    g.close()
