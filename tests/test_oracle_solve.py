"""CPU pins of the oracle's iterate-to-tolerance solve.

The solve has no counterpart in the reference (which runs a fixed
solve_step_count of Jacobi iterations, Simulation/Source/main.cpp:82), so it
is pinned by reducing it to the reference step: with tolerance 0 and I slots
it must be the plain step with I iterations, bit for bit; with a tolerance
above the first error it must be the step with no pressure update.  The warm
start is pinned by its two closed cases: zero multipliers change nothing,
and multipliers that push the error up are rejected.  Its retry at half the
multipliers reduces to the warm-started solve with the halved multipliers
one slot later, or to the plain solve two slots later.
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


def _c1_state(scenes):
    params, pos, vel = scenes.config_c1(seed=5)
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(pos, pos, vel)
    o.step(10)
    st = o.state()
    o.close()
    return params, st["x"].astype(np.float32), st["v"].astype(np.float32)


def _solve(params, x, v, *a, **kw):
    o = _oracle(params, x, v)
    used, hist, sums, kept = o.solve(*a, **kw)
    out = (used, hist, sums, o.last_solve["kept"], o.state(), o.counters()["hacks"])
    o.close()
    return out


def test_retry_applies_the_halved_start(scenes):
    """PBF_OPT_WARM_RETRY: a warm start that raises the mean error at slot 0
    is retried once at half its multipliers from the same position stars.
    When the half is kept, the solve is the warm-started solve with the halved
    multipliers one slot later: the same positions, velocities and sums, and
    its error history shifted by one.  Halving by a wrong factor (a quarter)
    or keeping the full start gives other histories."""
    params, x, v = _c1_state(scenes)
    _, _, lsum, _, _, _ = _solve(params, x, v, 0.0, 5)
    w = (lsum * np.float32(4.0)).astype(np.float32)
    half = (w * np.float32(0.5)).astype(np.float32)
    used, hist, sums, kept, st, hacks = _solve(params, x, v, 0.0, 4, warm=w, mean=True, retry=True)
    assert used == 4 and kept == 2
    used_n, hist_n, _, kept_n, _, _ = _solve(params, x, v, 0.0, 4, warm=w, mean=True)
    assert kept_n == 0 and hist_n[0] == hist[0]  # without the retry the start is dropped
    u2, h2, s2, k2, st2, hacks2 = _solve(params, x, v, 0.0, 3, warm=half, mean=True)
    assert k2 == 1 and hacks == hacks2
    assert np.array_equal(hist[1:], h2)
    assert np.array_equal(sums, s2)
    assert np.array_equal(st["x"], st2["x"]) and np.array_equal(st["v"], st2["v"])
    quarter = (w * np.float32(0.25)).astype(np.float32)
    _, h4, _, _, _, _ = _solve(params, x, v, 0.0, 3, warm=quarter, mean=True)
    assert not np.array_equal(hist[1:], h4)
    assert not np.array_equal(hist[1:], hist_n[1:])


def test_retry_falls_back_to_zero_when_the_half_grows_too(scenes):
    """Multipliers of the wrong sign grow the error at any scale: after the
    retry's slot the solve is the plain one, two slots later."""
    params, x, v = _scene(scenes, steps=8)
    w = np.full(len(x), 50.0, np.float32)
    used, hist, sums, kept, st, _ = _solve(params, x, v, 0.0, 4, warm=w, retry=True)
    assert kept == 0 and used == 4
    _, hp, sp, _, stp, _ = _solve(params, x, v, 0.0, 2)
    assert np.array_equal(hist[2:], hp)
    assert np.array_equal(sums, sp)
    assert np.array_equal(st["x"], stp["x"])


def test_retry_leaves_a_kept_start_alone(scenes):
    params, x, v = _c1_state(scenes)
    _, _, lsum, _, _, _ = _solve(params, x, v, 0.0, 5)
    a = _solve(params, x, v, 0.0, 4, warm=lsum, retry=True)
    b = _solve(params, x, v, 0.0, 4, warm=lsum)
    assert a[3] == 1 and b[3] == 1 and a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.array_equal(a[4]["x"], b[4]["x"])
