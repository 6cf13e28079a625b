"""Launch tuning changes no result; the autotuner restores the state it timed from."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(pbf, scenes, seed=2):
    params, pos, vel = scenes.config_c1(seed=seed)
    a = pbf.Particles(params)
    b = pbf.Particles(params)
    a.add(pos, pos, vel)
    b.add(pos, pos, vel)
    return a, b


@pytest.mark.parametrize("block,budget,hblocks", [(64, 64, 1), (128, 768, 8192), (256, 65536, 512)])
def test_tuning_changes_no_result(pbf, scenes, block, budget, hblocks):
    a, b = _pair(pbf, scenes)
    a.set_tuning(block, budget, hblocks)
    assert a.tuning() == {"block_size": block, "heavy_budget": budget, "heavy_blocks": hblocks}
    a.step(6)
    b.step(6)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.neighbor_lists()[0], b.neighbor_lists()[0])
    a.close(); b.close()


def test_autotune_restores_state(pbf, scenes):
    a, b = _pair(pbf, scenes, seed=7)
    a.step(3)
    b.step(3)
    best = a.autotune(steps=2)
    print(f"autotune: {best}", flush=True)
    assert best["block_size"] in (64, 128, 256) and best["ms_per_step"] > 0
    assert a.tuning() == {k: best[k] for k in ("block_size", "heavy_budget", "heavy_blocks")}
    a.step(4)
    b.step(4)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    a.close(); b.close()


def test_tuning_argument_errors(pbf, scenes):
    a, b = _pair(pbf, scenes)
    for args in ((96, 1536, 2048), (256, 10, 2048), (256, 1536, 0)):
        with pytest.raises(pbf.PBFError):
            a.set_tuning(*args)
    a.close(); b.close()
