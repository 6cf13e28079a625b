"""GPU parity at the full benchmark sizes, through the C ABI.

Three workloads against the CPU oracle (Jacobi mode, the device's semantics):
the 1,048,576-particle dam break (128 x 64 x 128 lattice, 200 substeps), the
8,388,608-particle double dam break (two 128x256x128 blocks, cap 64 with the
overflow counted, 100 substeps) and the batch of 256 dam breaks of
4,096..65,536 particles each, stepped by graph replay.

* The first substeps from the initial lattice: every field bit for bit.
* The state the whole run reaches, where the cap-60 collapse has piled
  particles into cells of hundreds of thousands: the neighbour lists the
  device builds for that state (the wavefront builder's far skips, the lane
  walk's adaptive budget) bit for bit, and for the dam break one more substep
  from it, which applies the near-coincidence fix tens of thousands of times.

For the collapsed states the oracle ends each list walk once its list is full
(Oracle.set_stop_at_cap): the same lists and counts as the full walk, which on
those cells takes minutes per pass.  The overflow count is compared where the
oracle walks in full (the first substep of the double dam break).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from test_gpu_parity import _compare_state

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
ROWS = 1 << 20  # rows of the lists compared at a time (bounds the host copies)


def _oracle(params, x, xs, v, stop_at_cap=False):
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_stop_at_cap(stop_at_cap)
    o.set_state(x, xs, v)
    return o


def _assert_lists_equal(g, o, cap, what):
    nb = o.neighbors()
    b, e = g.bin_bounds()
    assert np.array_equal(g.particle_ids(), nb["ids"]), what + ": particle order"
    assert np.array_equal(b, nb["begin"]) and np.array_equal(e, nb["end"]), what + ": cell bounds"
    lists, counts = g.neighbor_lists()
    assert np.array_equal(counts, nb["counts"]), what + ": counts"
    cols = np.arange(cap)[None, :]
    for r0 in range(0, counts.shape[0], ROWS):
        r1 = min(counts.shape[0], r0 + ROWS)
        mask = cols < counts[r0:r1, None]
        same = np.array_equal(np.where(mask, lists[r0:r1], 0), np.where(mask, nb["lists"][r0:r1], 0))
        assert same, f"{what}: lists of particles [{r0}, {r1})"
    return (e - b).max()


def test_dam_break_1m_first_substeps_bit_exact(pbf, scenes):
    p, pos, vel = scenes.config_c2(seed=0)
    p.options |= pbf.PBF_OPT_KEEP_WARM_FEATURES
    g = pbf.Particles(p)
    g.add(pos, pos, vel)
    o = _oracle(p, pos, pos, vel)
    g.step(2)
    o.step(2)
    _compare_state(g, o, "1M dam break, substep 2")


def test_dam_break_1m_collapsed_state_lists_and_next_substep(pbf, scenes):
    p, pos, vel = scenes.config_c2(seed=0)
    g = pbf.Particles(p)
    g.add(pos, pos, vel)
    g.step(200)
    x, xs, v = g.positions(), g.position_stars(), g.velocities()
    g.find_neighbors()
    o = _oracle(p, x, xs, v, stop_at_cap=True)
    o.pass_(O.FIND)
    biggest = _assert_lists_equal(g, o, p.max_neighbors, "1M dam break after 200 substeps")
    assert biggest > 100_000, f"largest cell holds {biggest} particles: the run did not collapse"
    o.set_state(x, x, v)
    g.step(1)
    o.step(1)
    assert o.counters()["hacks"] > 10_000
    _compare_state(g, o, "1M dam break, substep 201")


def test_double_dam_break_8m_first_substep_bit_exact(pbf, scenes):
    p, pos, vel = scenes.double_dam_break(seed=0, count_overflow=True)
    assert pos.shape[0] == 8_388_608 and p.max_neighbors == 64
    g = pbf.Particles(p)
    g.add(pos, pos, vel)
    o = _oracle(p, pos, pos, vel)
    g.step(1)
    o.step(1)
    _compare_state(g, o, "8M double dam break, substep 1")
    assert g.stats()["neighbor_overflow"] == o.counters()["overflow"]


def test_batched_256_scenes_graph_replay_bit_exact(pbf, scenes):
    """The whole 256-scene batch (4,096..65,536 particles per scene, log-uniform, about 5.7M in all),
    autotuned and stepped by graph replay as the benchmark does; six scenes (the smallest, the
    largest and four more) against the oracle run on their own, bit for bit."""
    params, boxes, sc = scenes.batched_dam_breaks(num_scenes=256, seed=0)
    g = pbf.Particles(params, boxes=boxes)
    for k, (pos, vel) in enumerate(sc):
        g.add_to_scene(k, pos, pos, vel)
    g.autotune(steps=2)  # restores the state it started from
    steps = 6
    g.step_graph(steps)
    x, v = g.positions(), g.velocities()
    sizes = np.array([pos.shape[0] for pos, _ in sc])
    first = np.concatenate([[0], np.cumsum(sizes)])
    assert first[-1] == x.shape[0] > 5_000_000
    picks = sorted({int(sizes.argmin()), int(sizes.argmax()), *np.random.default_rng(1).choice(256, 4).tolist()})
    for k in picks:
        pos, vel = sc[k]
        p = params.copy()
        p.max_particles_local = pos.shape[0]
        for i in range(3):
            p.boundary_max[i] = float(boxes[k, i])
        o = _oracle(p, pos, pos, vel)
        o.step(steps)
        st = o.state()
        a, b = first[k], first[k + 1]
        assert np.array_equal(x[a:b], st["x"].astype(np.float32)), f"scene {k} ({b - a} particles) positions"
        assert np.array_equal(v[a:b], st["v"].astype(np.float32)), f"scene {k} ({b - a} particles) velocities"
    g.close()


def test_double_dam_break_8m_full_run_lists_bit_exact(pbf, scenes):
    p, pos, vel = scenes.double_dam_break(seed=0, count_overflow=True)
    g = pbf.Particles(p)
    g.add(pos, pos, vel)
    g.step(100)
    x, xs, v = g.positions(), g.position_stars(), g.velocities()
    g.find_neighbors()
    o = _oracle(p, x, xs, v, stop_at_cap=True)
    o.pass_(O.FIND)
    _assert_lists_equal(g, o, p.max_neighbors, "8M double dam break after 100 substeps")


MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph-cpp_amd", "warm_model.npy")


def test_dam_break_1m_learned_warm_start(pbf, scenes):
    """The learned warm start on the 1,048,576-particle dam break with the
    trained network: after two substeps and one plain solve (whose multiplier
    sums are the previous sums), a PBF_WARM_LEARNED solve of four slots.  Its
    features must equal the oracle's bit for bit, and the whole solve must
    equal the oracle's solve warm-started with Lambda_prev + S * y, y being
    the network's output for those features (pbf_warm_model_predict): slots,
    error history, positions, velocities and multiplier sums."""
    if not os.path.exists(MODEL):
        pytest.skip("no trained model in the tree")
    p, pos, vel = scenes.config_c2(seed=0)
    p.options |= pbf.PBF_OPT_KEEP_WARM_FEATURES
    g = pbf.Particles(p)
    g.add(pos, pos, vel)
    o = _oracle(p, pos, pos, vel)
    g.step(2)
    o.step(2)
    _, _, lprev, _ = o.solve(0.0, 3)
    g.solve(0.0, 3)
    assert np.array_equal(g.lambda_sums(), lprev), "plain solve: multiplier sums"
    m = pbf.WarmModel(16)
    m.set_params(np.load(MODEL, allow_pickle=False))
    g.set_warm_model(m)
    print("1M learned warm start: device solve", flush=True)
    g.solve(0.0, 4, pbf.PBF_WARM_LEARNED)
    feat = g.warm_features()
    y = m.predict(feat)
    warm = (lprev + np.float32(pbf.PBF_WARM_LAMBDA_SCALE) * y).astype(np.float32)
    print("1M learned warm start: oracle solve", flush=True)
    used, hist, lsum, kept, ofeat = o.solve(0.0, 4, warm=warm, features=True)
    assert np.array_equal(feat, ofeat), "features"
    info = g.solve_info()
    print(f"1M learned warm start: slots {info['iterations']} (oracle {used}), kept {info['warm_start_kept']} "
          f"(oracle {kept}), nonzero outputs {int(np.count_nonzero(y))}", flush=True)
    assert info["iterations"] == used and bool(info["warm_start_kept"]) == kept
    assert np.array_equal(info["max_error_history"], hist)
    st = o.state()
    assert np.array_equal(g.positions(), st["x"].astype(np.float32)), "positions"
    assert np.array_equal(g.velocities(), st["v"].astype(np.float32)), "velocities"
    assert np.array_equal(g.lambda_sums(), lsum), "multiplier sums"
    g.close(); o.close(); m.close()


def test_dam_break_1m_network_gain_over_previous_sums(pbf, scenes):
    """What the trained network adds at 1M: the 1,048,576-particle dam break
    (neighbour cap 60), 200 substeps each iterated to a mean density error of
    1% (at most 128 slots), needs fewer Jacobi slots with the learned warm
    start than with the previous multiplier sums alone (PBF_WARM_PREVIOUS, the
    same start without the network's correction); measured 56.1 against 58.1
    slots per substep."""
    if not os.path.exists(MODEL):
        pytest.skip("no trained model in the tree")
    m = pbf.WarmModel(16)
    m.set_params(np.load(MODEL, allow_pickle=False))
    slots = {}
    for name, mode in (("previous", pbf.PBF_WARM_PREVIOUS), ("learned", pbf.PBF_WARM_LEARNED)):
        p, pos, vel = scenes.config_c2(seed=0)
        p.options |= pbf.PBF_OPT_SOLVE_MEAN_ERROR
        g = pbf.Particles(p)
        g.add(pos, pos, vel)
        if mode == pbf.PBF_WARM_LEARNED:
            g.set_warm_model(m)
        for st in range(0, 200, 50):
            print(f"1M network gain: {name}, substep {st}", flush=True)
            g.solve_graph(0.01, 128, mode, 50)
            g.sync()
        log = g.solve_log(200)
        slots[name] = sum(r["iterations"] for r in log)
        g.close()
    m.close()
    print(f"1M, 200 substeps: slots previous {slots['previous']}, learned {slots['learned']}", flush=True)
    assert slots["learned"] < slots["previous"]
