"""GPU parity of the iterate-to-tolerance solve and its warm starts.

The device runs the slot loop without host round trips (kernels of finished
slots return at once), so these tests check that the queued loop stops where
the oracle's host loop stops, bit for bit: slots used, the per-slot max
density error, positions, velocities and the multiplier sums.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _state(scenes, steps=10):
    """C1-sized dam break advanced `steps` substeps by the oracle (a dynamic state)."""
    params, pos, vel = scenes.config_c1(seed=5)
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(pos, pos, vel)
    o.step(steps)
    st = o.state()
    o.close()
    return params, st["x"].astype(np.float32), st["v"].astype(np.float32)


def _gpu(pbf, params, x, v):
    g = pbf.Particles(params)
    g.add(x, x, v)
    return g


def _orc(params, x, v):
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(x, x, v)
    return o


def _check(g, o, used, hist, lsum, kept, what):
    info = g.solve_info()
    print(f"{what}: slots {info['iterations']} (oracle {used}), kept {info['warm_start_kept']}", flush=True)
    assert info["iterations"] == used, what
    assert info["warm_start_kept"] == kept, what
    assert np.array_equal(info["max_error_history"], hist), what
    st = o.state()
    assert np.array_equal(g.positions(), st["x"].astype(np.float32)), what + ": positions"
    assert np.array_equal(g.velocities(), st["v"].astype(np.float32)), what + ": velocities"
    assert np.array_equal(g.lambda_sums(), lsum), what + ": multiplier sums"
    assert g.stats()["iterations"] == used


def test_solve_zero_tolerance_equals_fixed_step(pbf, scenes):
    params, x, v = _state(scenes)
    g = _gpu(pbf, params, x, v)
    g.solve(0.0, 4)
    h = _gpu(pbf, params, x, v)
    h.step(1)
    assert np.array_equal(g.positions(), h.positions())
    assert np.array_equal(g.velocities(), h.velocities())
    o = _orc(params, x, v)
    used, hist, lsum, kept = o.solve(0.0, 4)
    _check(g, o, used, hist, lsum, kept, "tol 0, 4 slots")
    g.close(); h.close(); o.close()


@pytest.mark.parametrize("k", [0, 3, 11])
def test_solve_stops_where_the_oracle_stops(pbf, scenes, k):
    params, x, v = _state(scenes)
    o = _orc(params, x, v)
    _, hist_all, _, _ = o.solve(0.0, 16)
    o.close()
    tol = float(hist_all[k])
    o = _orc(params, x, v)
    used, hist, lsum, kept = o.solve(tol, 16)
    g = _gpu(pbf, params, x, v)
    g.solve(tol, 16)
    _check(g, o, used, hist, lsum, kept, f"tol = e_{k}")
    g.close(); o.close()


def test_warm_start_given(pbf, scenes):
    params, x, v = _state(scenes)
    o = _orc(params, x, v)
    _, _, lam40, _ = o.solve(0.0, 40)
    o.close()
    o = _orc(params, x, v)
    used, hist, lsum, kept = o.solve(0.0, 6, warm=lam40)
    assert kept
    g = _gpu(pbf, params, x, v)
    g.set_warm_lambdas(lam40)
    g.solve(0.0, 6, pbf.PBF_WARM_GIVEN)
    info = g.solve_info()
    assert np.float32(info["max_error_before_warm_start"]) > info["max_error_history"][0]
    _check(g, o, used, hist, lsum, kept, "warm given")
    g.close(); o.close()


def test_warm_start_rejected(pbf, scenes):
    params, x, v = _state(scenes)
    bad = np.full(len(x), 50.0, np.float32)
    o = _orc(params, x, v)
    used, hist, lsum, kept = o.solve(0.0, 4, warm=bad)
    assert not kept
    g = _gpu(pbf, params, x, v)
    g.set_warm_lambdas(bad)
    g.solve(0.0, 4, pbf.PBF_WARM_GIVEN)
    _check(g, o, used, hist, lsum, kept, "warm rejected")
    g.close(); o.close()


def test_warm_start_previous_over_steps(pbf, scenes):
    """Three solves in a row, each warm-started from the previous one's sums
    (the particles are re-sorted between them on the device)."""
    params, x, v = _state(scenes)
    o = _orc(params, x, v)
    g = _gpu(pbf, params, x, v)
    used, hist, lsum, kept = o.solve(0.0, 5)
    g.solve(0.0, 5)
    _check(g, o, used, hist, lsum, kept, "previous, solve 0")
    for s in (1, 2):
        tol = float(hist[-1])  # stop once as good as the last plain slot
        used, hist, lsum, kept = o.solve(tol, 8, warm=lsum)
        g.solve(tol, 8, pbf.PBF_WARM_PREVIOUS)
        _check(g, o, used, hist, lsum, kept, f"previous, solve {s}")
    g.close(); o.close()


def test_solve_argument_errors(pbf, scenes):
    params, x, v = _state(scenes, steps=1)
    g = _gpu(pbf, params, x, v)
    with pytest.raises(pbf.PBFError):
        g.solve(0.0, 0)
    with pytest.raises(pbf.PBFError):
        g.solve(0.0, pbf.PBF_MAX_SOLVE_ITERATIONS + 1)
    with pytest.raises(pbf.PBFError):
        g.solve(-1.0, 4)
    with pytest.raises(pbf.PBFError):
        g.solve(0.0, 4, pbf.PBF_WARM_GIVEN)  # nothing given yet
    with pytest.raises(pbf.PBFError):
        g.solve_info()  # no solve has run
    g.close()


def test_solve_stops_at_a_slot_with_density_fixes(pbf, scenes):
    """The solve ends at a slot whose density pass moved particles (the
    near-coincidence fix, particles.h:303-305): every block of that slot's
    pressure update must still copy the moved x* (xt) to x*, including blocks
    that start after the first block has raised the done flag.  Every lattice
    point is doubled, so each density pass moves every particle of a pair;
    128,000 particles give 500 blocks of 256, more than the chip runs at once
    when each holds a full list walk."""
    lat = scenes.jitter(scenes.lattice((40, 40, 20)), seed=47)
    pos = np.concatenate([lat, lat]).astype(np.float32)
    params = scenes.pbf_params((2.0, 2.0, 1.0), pos.shape[0])
    v = np.zeros_like(pos)
    o = _orc(params, pos, v)
    _, hist_all, _, _ = o.solve(0.0, 8)
    o.close()
    # the first slot after slot 0 whose error is below every earlier one
    k = next(i for i in range(1, 8) if hist_all[i] < hist_all[:i].min())
    tol = float(hist_all[k])
    o = _orc(params, pos, v)
    used, hist, lsum, kept = o.solve(tol, 8)
    assert used == k + 1
    g = _gpu(pbf, params, pos, v)
    g.solve(tol, 8)
    assert g.stats()["density_hacks"] > 0
    _check(g, o, used, hist, lsum, kept, "stop at a slot with density fixes")
    g.close(); o.close()
