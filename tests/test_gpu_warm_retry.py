"""GPU parity of PBF_OPT_WARM_RETRY: a warm start that grows the error at slot
0 is retried once at half its multipliers before the solve falls back to a
zero warm start.  The oracle's solve (retry=True) is the checker: slots,
error histories, positions, velocities and multiplier sums bit for bit, for
a start whose half is kept, one whose half grows the error too, and the
previous-sum start over consecutive solves; graph replay and the learned
start are checked against the eager path and a given start.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _state(scenes, steps=10):
    params, pos, vel = scenes.config_c1(seed=5)
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(pos, pos, vel)
    o.step(steps)
    st = o.state()
    o.close()
    return params, st["x"].astype(np.float32), st["v"].astype(np.float32)


def _opts(pbf, params, mean=True, retry=True):
    p = pbf.Params.from_buffer_copy(params)
    if mean:
        p.options |= pbf.PBF_OPT_SOLVE_MEAN_ERROR
    if retry:
        p.options |= pbf.PBF_OPT_WARM_RETRY
    return p


def _orc(params, x, v):
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(x, x, v)
    return o


def _check(g, o, used, hist, lsum, mean, what):
    info = g.solve_info()
    print(f"{what}: slots {info['iterations']} (oracle {used}), kept {o.last_solve['kept']}", flush=True)
    assert info["iterations"] == used, what
    assert int(info["warm_start_kept"]) + int(info["warm_start_damped"]) == o.last_solve["kept"], what
    key = "mean_error_history" if mean else "max_error_history"
    assert np.array_equal(info[key], hist), what + ": error history"
    st = o.state()
    assert np.array_equal(g.positions(), st["x"].astype(np.float32)), what + ": positions"
    assert np.array_equal(g.velocities(), st["v"].astype(np.float32)), what + ": velocities"
    assert np.array_equal(g.lambda_sums(), lsum), what + ": multiplier sums"


@pytest.mark.parametrize("case", ["half kept", "half grows too"])
def test_retry_matches_oracle(pbf, scenes, case):
    params, x, v = _state(scenes)
    o = _orc(params, x, v)
    _, _, lsum5, _ = o.solve(0.0, 5)
    o.close()
    if case == "half kept":
        w = (lsum5 * np.float32(4.0)).astype(np.float32)  # grows the mean error; half of it does not
    else:
        w = np.full(len(x), 50.0, np.float32)  # the wrong sign: grows it at any scale
    o = _orc(params, x, v)
    used, hist, lsum, _ = o.solve(0.0, 6, warm=w, mean=True, retry=True)
    assert o.last_solve["kept"] == (2 if case == "half kept" else 0)
    g = pbf.Particles(_opts(pbf, params))
    g.add(x, x, v)
    g.set_warm_lambdas(w)
    g.solve(0.0, 6, pbf.PBF_WARM_GIVEN)
    _check(g, o, used, hist, lsum, True, case)
    g.close(); o.close()


def test_retry_with_previous_sums_over_solves(pbf, scenes):
    """Eight solves to a mean-error tolerance, each warm-started from the
    previous solve's sums, against the oracle fed the same sums."""
    params, x, v = _state(scenes, steps=14)
    o = _orc(params, x, v)
    _, hist0, _, _ = o.solve(0.0, 40, mean=True)
    o.close()
    tol = float(hist0[12])
    o = _orc(params, x, v)
    g = pbf.Particles(_opts(pbf, params))
    g.add(x, x, v)
    used, hist, lsum, _ = o.solve(tol, 64, mean=True)
    g.solve(tol, 64)
    _check(g, o, used, hist, lsum, True, "solve 0")
    kinds = []
    for k in range(1, 8):
        used, hist, lsum, _ = o.solve(tol, 64, warm=lsum, mean=True, retry=True)
        g.solve(tol, 64, pbf.PBF_WARM_PREVIOUS)
        kinds.append(o.last_solve["kept"])
        _check(g, o, used, hist, lsum, True, f"solve {k}")
    print("kept per solve (1 full, 2 halved, 0 none):", kinds, flush=True)
    g.close(); o.close()


def test_retry_graph_equals_eager(pbf, scenes):
    """pbf_solve_graph with the retry (the decision and the damped update
    happen on the device) against eager pbf_solve, step by step."""
    params, x, v = _state(scenes, steps=14)
    p = _opts(pbf, params)
    probe = pbf.Particles(p)
    probe.add(x, x, v)
    probe.solve(0.0, 40)
    tol = float(probe.solve_info()["mean_error_history"][12])
    probe.close()
    e = pbf.Particles(p)
    g = pbf.Particles(p)
    e.add(x, x, v)
    g.add(x, x, v)
    damped = 0
    for k in (2, 3, 2):
        for _ in range(k):
            e.solve(tol, 64, pbf.PBF_WARM_PREVIOUS)
            damped += int(e.solve_info()["warm_start_damped"])
        g.solve_graph(tol, 64, pbf.PBF_WARM_PREVIOUS, k)
        g.sync()
        for name in ("positions", "velocities", "lambda_sums"):
            assert np.array_equal(getattr(e, name)(), getattr(g, name)()), name
    log_e, log_g = e.solve_log(7), g.solve_log(7)
    assert [r["iterations"] for r in log_e] == [r["iterations"] for r in log_g]
    assert [r["warm_start_damped"] for r in log_e] == [r["warm_start_damped"] for r in log_g]
    print(f"damped starts kept: {damped} of 7", flush=True)
    e.close(); g.close()


def test_retry_learned_equals_given(pbf, scenes):
    """With the retry set, a learned solve equals a given solve fed the same
    prediction (the retry halves whichever multipliers the start used)."""
    params, x, v = _state(scenes)
    keep = _opts(pbf, params)
    keep.options |= pbf.PBF_OPT_KEEP_WARM_FEATURES
    m = pbf.WarmModel(len(x))
    m.init(7)
    a = pbf.Particles(keep)
    a.add(x, x, v)
    a.solve(0.0, 3)
    prev = a.lambda_sums()
    b = pbf.Particles(_opts(pbf, params))
    b.add(x, x, v)
    b.solve(0.0, 3)
    a.set_warm_model(m, gain=4.0)  # a large correction, so that the start is more likely to be retried
    a.solve(0.0, 6, pbf.PBF_WARM_LEARNED)
    warm = prev + np.float32(4.0 * pbf.PBF_WARM_LAMBDA_SCALE) * m.predict(a.warm_features())
    b.set_warm_lambdas(warm.astype(np.float32))
    b.solve(0.0, 6, pbf.PBF_WARM_GIVEN)
    ia, ib = a.solve_info(), b.solve_info()
    print(f"learned: slots {ia['iterations']}, kept {ia['warm_start_kept']}, damped {ia['warm_start_damped']}",
          flush=True)
    assert ia["iterations"] == ib["iterations"]
    assert ia["warm_start_damped"] == ib["warm_start_damped"]
    assert np.array_equal(ia["mean_error_history"], ib["mean_error_history"])
    for k in ("positions", "velocities", "lambda_sums"):
        assert np.array_equal(getattr(a, k)(), getattr(b, k)()), k
    a.close(); b.close(); m.close()
