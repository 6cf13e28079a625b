"""Batched scenes and graph-captured steps.

A batched context runs independent scenes in one set of launches: each
scene's cell keys are offset into its own key range, so its particles sort,
list and move exactly as in a context of their own.  The test compares every
scene of a batched run with its own single-scene run, bit for bit.  Graph
replays (pbf_step_graph) must equal eager steps bit for bit, also with eager
work between replays.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(scenes, k=5, seed=3):
    params, boxes, sc = scenes.batched_dam_breaks(num_scenes=k, seed=seed, nmin=1500, nmax=12000)
    return params, boxes, sc


def test_batched_equals_independent(pbf, scenes):
    params, boxes, sc = _scenes(scenes)
    g = pbf.Particles(params, boxes=boxes)
    assert g.num_scenes() == len(sc)
    for k, (pos, vel) in enumerate(sc):
        g.add_to_scene(k, pos, pos, vel)
    g.step(12)
    x = g.positions()
    v = g.velocities()
    keys_b = g.bin_ids()
    first = 0
    offset = 0
    print(f"batched: {len(sc)} scenes, {x.shape[0]} particles, {g.num_cells()} cells", flush=True)
    for k, (pos, vel) in enumerate(sc):
        p = params.copy()
        p.max_particles_local = pos.shape[0]
        for i in range(3):
            p.boundary_max[i] = float(boxes[k, i])
        h = pbf.Particles(p)
        h.add(pos, pos, vel)
        h.step(12)
        n = pos.shape[0]
        assert np.array_equal(h.positions(), x[first:first + n]), f"scene {k} positions"
        assert np.array_equal(h.velocities(), v[first:first + n]), f"scene {k} velocities"
        ncells = h.num_cells()
        # sorted keys of the batch: each scene's keys, shifted into its range
        assert np.array_equal(np.sort(keys_b[(keys_b >= offset) & (keys_b < offset + ncells)]) - offset,
                              np.sort(h.bin_ids()))
        first += n
        offset += ncells
        h.close()
    assert offset == g.num_cells()
    g.close()


def test_batched_scene_argument_errors(pbf, scenes):
    params, boxes, sc = _scenes(scenes, k=2)
    g = pbf.Particles(params, boxes=boxes)
    with pytest.raises(pbf.PBFError):
        g.add_to_scene(2, sc[0][0])
    g.close()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replay_equals_eager(pbf, scenes, batched):
    if batched:
        params, boxes, sc = _scenes(scenes, k=4, seed=8)
        a = pbf.Particles(params, boxes=boxes)
        b = pbf.Particles(params, boxes=boxes)
        for k, (pos, vel) in enumerate(sc):
            a.add_to_scene(k, pos, pos, vel)
            b.add_to_scene(k, pos, pos, vel)
    else:
        params, pos, vel = scenes.config_c1(seed=4)
        a = pbf.Particles(params)
        b = pbf.Particles(params)
        a.add(pos, pos, vel)
        b.add(pos, pos, vel)
    # replays, eager steps in between (odd counts end eagerly), several checks
    for n in (4, 1, 3, 2, 6):
        a.step_graph(n)
        b.step(n)
        a.sync()
        assert np.array_equal(a.positions(), b.positions()), f"after {n} more steps"
        assert np.array_equal(a.velocities(), b.velocities())
    # a parameter change recaptures
    p2 = params.copy()
    p2.visc_c = np.float32(0.02)
    a.set_params(p2)
    b.set_params(p2)
    a.step_graph(4)
    b.step(4)
    assert np.array_equal(a.positions(), b.positions())
    a.close()
    b.close()


def test_graph_replay_with_density_fixes(pbf, scenes):
    """Replays of a scene where the near-coincidence fix (particles.h:303-305)
    fires thousands of times per step (every lattice point doubled): the fix's
    moved and dirty stamps repeat from replay to replay and are cleared inside
    the graph; replays must still equal eager steps bit for bit."""
    lat = scenes.jitter(scenes.lattice((16, 16, 16)), seed=43)
    pos = np.concatenate([lat, lat]).astype(np.float32)
    params = scenes.pbf_params((1.6, 1.6, 1.6), pos.shape[0])
    a = pbf.Particles(params)
    b = pbf.Particles(params)
    a.add(pos, pos, np.zeros_like(pos))
    b.add(pos, pos, np.zeros_like(pos))
    hacks = 0
    for n in (2, 4, 2):
        a.step_graph(n)
        b.step(n)
        a.sync()
        hacks += b.stats()["density_hacks"]
        assert np.array_equal(a.positions(), b.positions()), f"after {n} more steps"
        assert np.array_equal(a.velocities(), b.velocities())
    assert hacks > 0
    a.close()
    b.close()


def test_batched_density_fixes_equal_independent(pbf, scenes):
    """The near-coincidence fix (particles.h:303-305) inside a batched context:
    scene 0 has one coincident pair (neighbourhood marking through the
    scene's own grid), scene 1 doubles every lattice point (more than 4,096
    moved particles: the exact scan).  Each scene must evolve exactly as in a
    context of its own."""
    a = scenes.jitter(scenes.lattice((12, 12, 12)), seed=51)
    a[10] = a[11]
    lat = scenes.jitter(scenes.lattice((16, 16, 16)), seed=52)
    b = np.concatenate([lat, lat]).astype(np.float32)
    boxes = np.array([[1.2, 1.2, 1.2], [1.6, 1.6, 1.6]], dtype=np.float32)
    parts = ((a, boxes[0]), (b, boxes[1]))
    params = scenes.pbf_params((1.6, 1.6, 1.6), a.shape[0] + b.shape[0])
    g = pbf.Particles(params, boxes=boxes)
    for k, (pos, _) in enumerate(parts):
        g.add_to_scene(k, pos, pos, np.zeros_like(pos))
    g.step(3)
    x, v = g.positions(), g.velocities()
    first = 0
    for k, (pos, box) in enumerate(parts):
        p = params.copy()
        p.max_particles_local = pos.shape[0]
        for i in range(3):
            p.boundary_max[i] = float(box[i])
        h = pbf.Particles(p)
        h.add(pos, pos, np.zeros_like(pos))
        h.step(3)
        n = pos.shape[0]
        print(f"scene {k}: {n} particles, last-step density fixes {h.stats()['density_hacks']}", flush=True)
        assert h.stats()["density_hacks"] > 0 or k == 0
        assert np.array_equal(h.positions(), x[first:first + n]), f"scene {k} positions"
        assert np.array_equal(h.velocities(), v[first:first + n]), f"scene {k} velocities"
        first += n
        h.close()
    g.close()
