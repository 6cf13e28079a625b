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


def test_warm_start_strict_fallback(pbf, scenes):
    """PBF_OPT_WARM_STRICT: the warm start is also rejected when the error
    measure the tolerance does not use grows.  A given start of 4x the sums of
    a plain 5-slot solve lowers the max error but raises the mean: kept by the
    default fallback, rejected by the strict one; both bit for bit against the
    oracle."""
    params, x, v = _state(scenes)
    o = _orc(params, x, v)
    _, _, lsum, _ = o.solve(0.0, 5)
    o.close()
    w = (lsum * np.float32(4.0)).astype(np.float32)
    kept_by = {}
    for strict in (False, True):
        p = pbf.Params.from_buffer_copy(params)
        if strict:
            p.options |= pbf.PBF_OPT_WARM_STRICT
        o = _orc(params, x, v)
        used, hist, lsum2, kept = o.solve(0.0, 4, warm=w, strict=strict)
        g = _gpu(pbf, p, x, v)
        g.set_warm_lambdas(w)
        g.solve(0.0, 4, pbf.PBF_WARM_GIVEN)
        _check(g, o, used, hist, lsum2, kept, f"given 4x sums, strict {strict}")
        kept_by[strict] = kept
        g.close(); o.close()
    assert kept_by == {False: True, True: False}


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


def _solve_state(g):
    return g.positions(), g.velocities(), g.lambda_sums()


@pytest.mark.parametrize("warm", ["none", "previous", "learned"])
def test_solve_graph_equals_eager_solve(pbf, scenes, warm):
    """pbf_solve_graph (two captured solves replayed, odd counts finished by
    one queued solve) against pbf_solve, step by step: positions, velocities,
    multiplier sums and the per-solve records, bit for bit.  The graph is
    replayed several times with eager solves and host reads between the
    replays, and every replay is checked."""
    mode = {"none": pbf.PBF_WARM_NONE, "previous": pbf.PBF_WARM_PREVIOUS, "learned": pbf.PBF_WARM_LEARNED}[warm]
    params, x, v = _state(scenes)
    model = None
    if mode == pbf.PBF_WARM_LEARNED:
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-cpp_amd",
                            "warm_model.npy")
        model = pbf.WarmModel(16)
        model.set_params(np.load(path, allow_pickle=False))
    e = _gpu(pbf, params, x, v)
    g = _gpu(pbf, params, x, v)
    if model is not None:
        e.set_warm_model(model)
        g.set_warm_model(model)
    # a tolerance the plain solve of the first step meets at its 11th slot
    probe = _gpu(pbf, params, x, v)
    probe.solve(0.0, 24)
    tol, slots = float(probe.solve_info()["max_error_history"][10]), 24
    probe.close()
    infos = []
    # graph calls of 2, 3 (2 + 1 queued), 4 and 2 steps, with eager solves between
    plan = [("graph", 2), ("eager", 1), ("graph", 3), ("graph", 4), ("eager", 1), ("graph", 2)]
    done = 0
    for kind, k in plan:
        for _ in range(k):
            e.solve(tol, slots, mode)
            infos.append(e.solve_info())
        if kind == "graph":
            g.solve_graph(tol, slots, mode, k)
        else:
            g.solve(tol, slots, mode)
        done += k
        what = f"{warm}: after {done} steps ({kind} {k})"
        print(what, flush=True)
        for a, b, name in zip(_solve_state(g), _solve_state(e), ("positions", "velocities", "sums")):
            assert np.array_equal(a, b), what + ": " + name
    log = g.solve_log(done)
    assert [r["index"] for r in log] == list(range(done))
    for r, info in zip(log, infos):
        k = info["iterations"]
        assert r["iterations"] == k
        assert r["warm_start_kept"] == info["warm_start_kept"]
        assert np.float32(r["max_error"]) == info["max_error_history"][k - 1]
        assert np.float32(r["mean_error"]) == info["mean_error_history"][k - 1]
    assert any(r["iterations"] < slots for r in log), "no solve ended before its last slot"
    gi = g.solve_info()
    assert gi["iterations"] == infos[-1]["iterations"]
    assert np.array_equal(gi["max_error_history"], infos[-1]["max_error_history"])
    e.close(); g.close()
    if model is not None:
        model.close()


def test_solve_graph_recaptures_on_new_arguments(pbf, scenes):
    """A new tolerance or slot limit gives a new graph: the results follow the
    arguments of each call (checked against eager solves)."""
    params, x, v = _state(scenes)
    e = _gpu(pbf, params, x, v)
    g = _gpu(pbf, params, x, v)
    for tol, slots in ((0.05, 8), (0.02, 16), (0.02, 5), (0.05, 8)):
        for _ in range(2):
            e.solve(tol, slots)
        g.solve_graph(tol, slots, pbf.PBF_WARM_NONE, 2)
        for a, b in zip(_solve_state(g), _solve_state(e)):
            assert np.array_equal(a, b), (tol, slots)
    e.close(); g.close()


def test_solve_graph_argument_errors(pbf, scenes):
    params, x, v = _state(scenes, steps=0)
    g = _gpu(pbf, params, x, v)
    for args in ((-1.0, 4, 0, 2), (0.01, 0, 0, 2), (0.01, pbf.PBF_MAX_SOLVE_ITERATIONS + 1, 0, 2),
                 (0.01, 4, 9, 2)):
        with pytest.raises(pbf.PBFError):
            g.solve_graph(*args)
    with pytest.raises(pbf.PBFError):
        g.solve_graph(0.01, 4, pbf.PBF_WARM_LEARNED, 2)  # no model
    g.close()
