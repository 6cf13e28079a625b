"""CPU pins of the oracle's iterate-to-tolerance solve.

The solve has no counterpart in the reference (which runs a fixed
solve_step_count of Jacobi iterations, Simulation/Source/main.cpp:82), so it
is pinned by reducing it to the reference step: with tolerance 0 and I slots
it must be the plain step with I iterations, bit for bit; with a tolerance
above the first error it must be the step with no pressure update.  The warm
start is pinned by its two closed cases: zero multipliers change nothing,
and multipliers that push the error up are rejected.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

THREADS = max(1, min(8, os.cpu_count() or 1))


def _scene(scenes, steps=0, iters=4):
    params, pos, vel = scenes.dam_break((12, 10, 8), (1.2, 1.6, 0.8), seed=3, iters=iters)
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(pos, pos, vel)
    if steps:
        o.step(steps)
    st = o.state()
    x, v = st["x"].astype(np.float32), st["v"].astype(np.float32)
    o.close()
    return params, x, v


def _oracle(params, x, v, iters=None):
    op = O.params_from_pbf(params)
    if iters is not None:
        op.iters = iters
    o = O.Oracle(op, mode=O.JACOBI, threads=THREADS)
    o.set_state(x, x, v)
    return o


def _same(a, b):
    sa, sb = a.state(), b.state()
    for k in ("x", "v"):
        assert np.array_equal(sa[k], sb[k]), k


def test_zero_tolerance_is_the_fixed_iteration_step(scenes):
    params, x, v = _scene(scenes, steps=8)
    a = _oracle(params, x, v, iters=4)
    a.step(1)
    b = _oracle(params, x, v)
    used, hist, lsum, kept = b.solve(0.0, 4)
    assert used == 4 and len(hist) == 4 and not kept
    assert np.all(hist > 0)
    _same(a, b)


def test_large_tolerance_stops_in_the_first_slot(scenes):
    params, x, v = _scene(scenes, steps=8)
    # the step body with one density pass and no pressure update
    a = _oracle(params, x, v)
    for pid in (0, 1, 2, 3, 7, 9, 10, 11, 9, 12):
        a.pass_(pid)
    b = _oracle(params, x, v)
    used, hist, lsum, kept = b.solve(1e9, 10)
    assert used == 1
    assert b.counters()["hacks"] == 0  # otherwise the slot-0 density fix would separate a from b
    _same(a, b)
    assert np.all(lsum == 0)


def test_slot_count_follows_the_tolerance(scenes):
    params, x, v = _scene(scenes, steps=8)
    o = _oracle(params, x, v)
    used, hist, _, _ = o.solve(0.0, 12)
    o.close()
    assert used == 12
    # the first slot whose error is <= tol ends the solve
    for k in (2, 5, 9):
        tol = float(hist[k])
        expect = int(np.argmax(hist <= np.float32(tol))) + 1
        o = _oracle(params, x, v)
        used_k, hist_k, _, _ = o.solve(tol, 12)
        o.close()
        assert used_k == expect
        assert np.array_equal(hist_k, hist[:expect])


def test_lambda_sum_is_the_sum_of_applied_multipliers(scenes):
    params, x, v = _scene(scenes, steps=8)
    o = _oracle(params, x, v)
    used, hist, lsum, _ = o.solve(0.0, 3)
    # closed form: replay the same three iterations pass by pass
    r = _oracle(params, x, v)
    for pid in (0, 1, 2):
        r.pass_(pid)
    acc = np.zeros(len(x), np.float32)
    for _ in range(3):
        r.pass_(3)
        r.pass_(4)
        acc += r.state()["lam"].astype(np.float32)
        r.pass_(5)
        r.pass_(6)
    assert np.array_equal(lsum, acc)
    assert np.all(lsum <= 0)  # lambda = -C / (...) with C >= 0


def test_zero_warm_start_changes_nothing(scenes):
    params, x, v = _scene(scenes, steps=8)
    a = _oracle(params, x, v)
    ua, ha, la, _ = a.solve(0.0, 5)
    assert a.counters()["hacks"] == 0
    b = _oracle(params, x, v)
    ub, hb, lb, kept = b.solve(0.0, 5, warm=np.zeros(len(x), np.float32))
    assert kept and ua == ub
    assert np.array_equal(ha, hb) and np.array_equal(la, lb)
    _same(a, b)


def test_harmful_warm_start_is_rejected(scenes):
    params, x, v = _scene(scenes, steps=8)
    # multipliers of the wrong sign pull neighbours together: the error grows
    used, hist, lsum, kept = _oracle(params, x, v).solve(0.0, 3, warm=np.full(len(x), 50.0, np.float32))
    assert not kept
    # after the rejected slot the solve is the plain one, one slot later
    a = _oracle(params, x, v)
    ua, ha, la, _ = a.solve(0.0, 2)
    assert np.array_equal(hist[1:], ha)
    assert np.array_equal(lsum, la)


def test_converged_warm_start_lowers_the_error(scenes):
    params, x, v = _scene(scenes, steps=8)
    _, h0, lsum, _ = _oracle(params, x, v).solve(0.0, 40)
    _, hw, _, kept = _oracle(params, x, v).solve(0.0, 4, warm=lsum)
    assert kept
    assert hw[0] < h0[0]


def test_strict_fallback_rejects_what_raises_either_error(scenes):
    """The strict fallback (PBF_OPT_WARM_STRICT) rejects a warm start that
    lowers the max error but raises the mean; the default keeps it.  A
    rejected start leaves the plain solve with its first slot spent: its
    positions and sums equal the plain solve with one slot fewer (no
    near-coincidence fix fires in this state)."""
    params, pos, vel = scenes.config_c1(seed=5)
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(pos, pos, vel)
    o.step(10)
    st = o.state()
    o.close()
    x, v = st["x"].astype(np.float32), st["v"].astype(np.float32)
    a = _oracle(params, x, v)
    _, _, lsum, _ = a.solve(0.0, 5)
    a.close()
    w = (lsum * np.float32(4.0)).astype(np.float32)
    res = {}
    for strict in (False, True):
        b = _oracle(params, x, v)
        used, hist, sums, kept = b.solve(0.0, 4, warm=w, strict=strict)
        mx, mn = b.last_solve["max"], b.last_solve["mean"]
        res[strict] = (kept, b.state(), sums, b.counters()["hacks"])
        b.close()
    assert res[False][0] is True and res[True][0] is False
    plain = _oracle(params, x, v)
    _, _, psums, _ = plain.solve(0.0, 3)
    pst = plain.state()
    plain.close()
    assert res[True][3] == 0
    assert np.array_equal(res[True][1]["x"], pst["x"])
    assert np.array_equal(res[True][2], psums)
