"""Non-finite state and the mover's singular point on the device.

The behaviour is stated and pinned on the oracle in tests/test_oracle_failures.py:
a particle at the mover's centre, or with a NaN position or velocity, gets a
NaN x* (particles.h:519-521), sits in the outside bucket that no walk visits,
and leaves every other particle exactly as without it; an infinite velocity
is clamped onto the box face.  Here the library must give the same state as
the oracle, bit for bit (NaN where the oracle has NaN).
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _scene(scenes):
    return scenes.dam_break((8, 8, 8), (0.8, 0.8, 0.8), seed=21)


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def _run(pbf, params, pos, vel, steps):
    g = pbf.Particles(params)
    g.add(pos, pos, vel)
    g.step(steps)
    out = {"x": g.positions(), "v": g.velocities(), "rho": g.densities(), "lam": g.lambdas(),
           "outside": g.stats()["outside_grid"]}
    g.close()
    return out


def _oracle(params, pos, vel, steps):
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=1)
    o.set_state(pos, pos, vel)
    o.step(steps)
    st = o.state()
    out = {k: st[k].astype(np.float32) for k in ("x", "v", "rho", "lam")}
    out["outside"] = o.counters()["outside"]
    o.close()
    return out


def _compare(got, exp, what):
    for f in ("x", "v", "rho", "lam"):
        assert _same(got[f], exp[f]), f"{what}: {f}"
    assert got["outside"] == exp["outside"], what


def test_mover_exactly_on_a_particle(pbf, scenes):
    params, pos, vel = _scene(scenes)
    i = 200
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=1)
    o.set_state(pos, pos, vel)
    o.pass_(O.FORCES)
    o.pass_(O.PREDICT)
    c = o.state()["xs"][i].astype(np.float32)
    o.close()
    p = params.copy()
    for k in range(3):
        p.mover_center[k] = float(c[k])
    print(f"mover at particle {i}: {c}", flush=True)
    got, exp = _run(pbf, p, pos, vel, 3), _oracle(p, pos, vel, 3)
    assert np.isnan(got["x"][i]).all()
    _compare(got, exp, "mover on a particle")


@pytest.mark.parametrize("where", ["position", "velocity"])
def test_nan_in_the_state(pbf, scenes, where):
    params, pos, vel = _scene(scenes)
    pos, vel = pos.copy(), vel.copy()
    (pos if where == "position" else vel)[77, 1] = np.nan
    print(f"NaN in the {where} of particle 77", flush=True)
    got, exp = _run(pbf, params, pos, vel, 2), _oracle(params, pos, vel, 2)
    _compare(got, exp, f"NaN {where}")


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_infinite_velocity(pbf, scenes, sign):
    params, pos, vel = _scene(scenes)
    vel = vel.copy()
    vel[5, 0] = sign * np.inf
    print(f"velocity {sign} inf on particle 5", flush=True)
    got, exp = _run(pbf, params, pos, vel, 2), _oracle(params, pos, vel, 2)
    assert np.isfinite(got["x"]).all() and np.isfinite(got["v"]).all()
    _compare(got, exp, "infinite velocity")
