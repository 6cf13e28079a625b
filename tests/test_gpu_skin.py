"""GPU tests of the skin lists: the fused step's Jacobi passes (and its first
XSPH pass) walk each list's entries within 1.1 h of the positions the lists
were built from, in list order, while no particle has moved 0.045 h since;
past that move, or when the previous step's moves went past it, they walk the
full lists.  Either way the sums run over the same entries in the same order,
so a context with skin lists and one with PBF_OPT_FULL_LISTS must stay
identical bit for bit, through the calm start of a dam break, its collapse
(where the moves pass the limit and the lists switch off and back on) and the
iterate-to-tolerance solves with warm starts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(pbf, params, pos, vel, boxes=None, scenes_list=None):
    out = []
    for full in (False, True):
        p = pbf.Params.from_buffer_copy(params)
        if full:
            p.options |= pbf.PBF_OPT_FULL_LISTS
        g = pbf.Particles(p, boxes=boxes)
        if scenes_list is None:
            g.add(pos, pos, vel)
        else:
            for k, (x, v) in enumerate(scenes_list):
                g.add_to_scene(k, x, x, v)
        out.append(g)
    return out


def _same(a, b, what):
    assert np.array_equal(a.positions(), b.positions()), what + ": positions"
    assert np.array_equal(a.velocities(), b.velocities()), what + ": velocities"


def test_skin_lists_change_no_result_over_a_dam_break(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=2)
    s, f = _pair(pbf, params, pos, vel)
    for k in range(10):
        s.step(10)
        f.step(10)
        _same(s, f, f"substep {10 * (k + 1)}")
    s.close(); f.close()


def test_skin_lists_change_no_result_in_solves(pbf, scenes):
    params, pos, vel = scenes.config_c1(seed=3)
    params.options |= pbf.PBF_OPT_SOLVE_MEAN_ERROR
    s, f = _pair(pbf, params, pos, vel)
    for mode in (pbf.PBF_WARM_NONE, pbf.PBF_WARM_PREVIOUS):
        s.solve_graph(0.01, 64, mode, 15)
        f.solve_graph(0.01, 64, mode, 15)
        _same(s, f, f"solves, warm start {mode}")
        assert np.array_equal(s.lambda_sums(), f.lambda_sums())
        ls, lf = s.solve_log(15), f.solve_log(15)
        assert [r["iterations"] for r in ls] == [r["iterations"] for r in lf]
    s.close(); f.close()


def test_skin_lists_change_no_result_in_batched_graph_replay(pbf, scenes):
    params, boxes, sc = scenes.batched_dam_breaks(num_scenes=8, seed=4, nmin=2048, nmax=16384)
    s, f = _pair(pbf, params, None, None, boxes=boxes, scenes_list=sc)
    s.step_graph(6)
    f.step_graph(6)
    _same(s, f, "batched, 6 graph-replayed steps")
    s.close(); f.close()
