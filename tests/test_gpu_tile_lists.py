"""GPU tests of the cell-tile list walk, the list builder's default for the
particles of ordinary cells, against the per-lane walk (PBF_OPT_LANE_LISTS):
lists, counts and overflow counts must be the same bit for bit, and so must
the steps that walk them (through the skin lists when those are used).  The
oracle comparisons in tests/test_gpu_parity.py run the default builder too;
these tests also check, through the tile_listed statistic, that the tile walk
is the one that built most lists.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _params(pbf, params, lane):
    p = pbf.Params.from_buffer_copy(params)
    if lane:
        p.options |= pbf.PBF_OPT_LANE_LISTS
    return p


def _pair(pbf, params, pos=None, vel=None, boxes=None, scenes_list=None):
    out = []
    for lane in (False, True):
        g = pbf.Particles(_params(pbf, params, lane), boxes=boxes)
        if scenes_list is None:
            g.add(pos, pos, vel)
        else:
            for k, (x, v) in enumerate(scenes_list):
                g.add_to_scene(k, x, x, v)
        out.append(g)
    return out


def _lists_equal(a, b, what):
    la, ca = a.neighbor_lists()
    lb, cb = b.neighbor_lists()
    assert np.array_equal(ca, cb), what + ": counts"
    mask = np.arange(la.shape[1])[None, :] < ca[:, None]
    assert np.array_equal(np.where(mask, la, 0), np.where(mask, lb, 0)), what + ": lists"


def _state_equal(a, b, what):
    assert np.array_equal(a.positions(), b.positions()), what + ": positions"
    assert np.array_equal(a.velocities(), b.velocities()), what + ": velocities"


def _counting(pbf, params):
    p = pbf.Params.from_buffer_copy(params)
    p.options |= pbf.PBF_OPT_COUNT_OVERFLOW
    return p


def test_tile_walk_over_a_dam_break(pbf, scenes):
    """The 32,768-particle dam break, 100 substeps in lockstep: the lists after
    every tenth substep and the state after it.  The skin lists are built by
    the same walk, so the steps in between check them as well."""
    params, pos, vel = scenes.config_c1(seed=8)
    t, l = _pair(pbf, _counting(pbf, params), pos, vel)
    t.find_neighbors()
    l.find_neighbors()
    _lists_equal(t, l, "initial lists")
    st, sl = t.stats(), l.stats()
    print(f"tile walk: {st['tile_listed']} of {pos.shape[0]} particles; lane walk option: {sl['tile_listed']}",
          flush=True)
    assert st["tile_listed"] >= 0.9 * pos.shape[0]
    assert sl["tile_listed"] == 0
    assert st["neighbor_overflow"] == sl["neighbor_overflow"]
    for k in range(10):
        t.step(10)
        l.step(10)
        _state_equal(t, l, f"substep {10 * (k + 1)}")
        _lists_equal(t, l, f"lists after substep {10 * (k + 1)}")
        assert t.stats()["neighbor_overflow"] == l.stats()["neighbor_overflow"]
    t.close(); l.close()


@pytest.mark.parametrize("cap", [24, 60, 128])
def test_tile_walk_caps_and_overflow(pbf, scenes, cap):
    """A lattice at 0.4 h spacing (about 130 neighbours within the bin
    spacing): the cap cuts every list (24, 60) or none (128); lists and the
    overflow count must match the lane walk's."""
    pos = scenes.jitter(scenes.lattice((12, 10, 10), spacing=0.04), seed=17)
    p = scenes.pbf_params((1.2, 0.8, 0.8), pos.shape[0], cap=cap, count_overflow=True)
    t, l = _pair(pbf, p, pos, np.zeros_like(pos))
    t.find_neighbors()
    l.find_neighbors()
    _lists_equal(t, l, f"cap {cap}")
    st, sl = t.stats(), l.stats()
    assert st["neighbor_overflow"] == sl["neighbor_overflow"]
    assert (st["neighbor_overflow"] > 0) == (cap < 128)
    assert st["tile_listed"] >= 0.5 * pos.shape[0]
    t.step(2)
    l.step(2)
    _state_equal(t, l, f"cap {cap}, 2 steps")
    t.close(); l.close()


def test_tile_walk_mixed_with_crowded_cells(pbf, scenes):
    """A dense cluster (cells past the tile walk's candidate limit, listed by
    the lane and wavefront walks) below an ordinary lattice (cell layers of
    their own in the z-major cell order, listed by the tile walk)."""
    rng = np.random.default_rng(21)
    dense = rng.uniform(0.30, 0.40, size=(3000, 3)).astype(np.float32)
    sparse = scenes.lattice((16, 16, 4), origin=(0.05, 0.05, 0.6))
    pos = np.concatenate([dense, sparse]).astype(np.float32)
    p = scenes.pbf_params((1.0, 1.0, 1.0), pos.shape[0], count_overflow=True)
    t, l = _pair(pbf, p, pos, np.zeros_like(pos))
    t.find_neighbors()
    l.find_neighbors()
    _lists_equal(t, l, "clusters and lattice")
    st = t.stats()
    assert st["neighbor_overflow"] == l.stats()["neighbor_overflow"] > 0
    assert 0 < st["tile_listed"] < pos.shape[0]
    t.close(); l.close()


def _far_face_params(pbf, n):
    # neighbor_test.ini's box: 5 x 5 x 5, bin spacing 1.0 (an exact divisor of the extent)
    p = pbf.Params()
    p.max_particles_local = n
    p.solve_step_count = 4
    p.max_neighbors = 60
    p.particle_rest_spacing = 0.5
    p.smoothing_radius = 1.0
    p.neighbor_bin_spacing = 1.0
    p.time_step = 1.0 / 60.0
    for i in range(3):
        p.boundary_max[i] = 5.0
        p.mover_center[i] = -10.0
    p.options = pbf.PBF_OPT_COUNT_OVERFLOW
    pbf.derive(p)
    return p


def test_tile_walk_at_the_far_faces(pbf, scenes):
    """Particles on the box's far faces (x, y or z = 5.0 with bin spacing 1.0)
    sit in the last cell column, so one of their neighbour columns is one past
    the grid: the reference's key arithmetic wraps it into the next row (or
    plane), and both walks must follow it.  Lists against the lane walk and
    the oracle."""
    rng = np.random.default_rng(12)
    inner = rng.uniform(0.0, 5.0, size=(900, 3)).astype(np.float32)
    face = rng.uniform(0.0, 5.0, size=(300, 3)).astype(np.float32)
    face[:100, 0] = 5.0
    face[100:200, 1] = 5.0
    face[200:, 2] = 5.0
    pos = np.concatenate([inner, face]).astype(np.float32)
    p = _far_face_params(pbf, pos.shape[0])
    t, l = _pair(pbf, p, pos, np.zeros_like(pos))
    print(f"grid dimensions {t.bin_dimensions()}", flush=True)
    t.find_neighbors()
    l.find_neighbors()
    _lists_equal(t, l, "far faces")
    assert t.stats()["neighbor_overflow"] == l.stats()["neighbor_overflow"]
    assert t.stats()["tile_listed"] > 0
    o = O.Oracle(O.params_from_pbf(p), mode=O.JACOBI, threads=1)
    o.set_state(pos, pos, np.zeros_like(pos))
    o.pass_(O.FIND)
    nb = o.neighbors()
    lists, counts = t.neighbor_lists()
    assert np.array_equal(counts, nb["counts"])
    mask = np.arange(p.max_neighbors)[None, :] < counts[:, None]
    assert np.array_equal(np.where(mask, lists, 0), np.where(mask, nb["lists"], 0))
    o.close()
    t.close(); l.close()


def test_tile_walk_batched_scenes(pbf, scenes):
    params, boxes, sc = scenes.batched_dam_breaks(num_scenes=8, seed=4, nmin=2048, nmax=16384)
    t, l = _pair(pbf, _counting(pbf, params), boxes=boxes, scenes_list=sc)
    t.find_neighbors()
    l.find_neighbors()
    _lists_equal(t, l, "batched scenes")
    assert t.stats()["tile_listed"] >= 0.8 * sum(x.shape[0] for x, _ in sc)
    t.step_graph(6)
    l.step_graph(6)
    _state_equal(t, l, "batched, 6 graph-replayed steps")
    t.close(); l.close()


def test_tile_walk_million_particles(pbf, scenes):
    """The 1,048,576-particle dam break: lists at the start, and 20 substeps."""
    params, pos, vel = scenes.config_c2(seed=0)
    t, l = _pair(pbf, _counting(pbf, params), pos, vel)
    t.find_neighbors()
    l.find_neighbors()
    _lists_equal(t, l, "1M lists")
    st = t.stats()
    print(f"1M: tile walk listed {st['tile_listed']} particles", flush=True)
    assert st["tile_listed"] >= 0.9 * pos.shape[0]
    t.step(20)
    l.step(20)
    _state_equal(t, l, "1M, 20 substeps")
    t.close(); l.close()
