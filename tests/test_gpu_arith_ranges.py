"""GPU tests of the two instances of the exact pressure-solve arithmetic.

The correctly rounded divide and square root in the density and pressure
passes carry range tests (operands where their fast sequences would differ
from the plain operations).  For the pairs the within-h sums accept, the host
discharges most of them when the spiky numerators |spiky_norm| h^3 stay at or
below 2^40, i.e. for smoothing radii from about 2.4e-4 up; smaller radii keep
the full tests.  Scaled copies of a dam break on both sides of that line (and
one just inside it) step and solve bit for bit with the oracle.
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _static_ranges(h):
    """The host's condition on the numerators (pbf_api.hip, compute_kparams)."""
    h = np.float32(h)
    spiky = np.float32(-45.0 / (math.pi * float(h) ** 6))
    return abs(float(spiky)) * float(h) ** 3 * 1.001 <= 2.0 ** 40


def _scene(scenes, scale, seed):
    h, s = 0.1 * scale, 0.05 * scale
    pos = scenes.jitter(scenes.lattice((16, 16, 16), spacing=s), seed=seed, amplitude=0.01 * h)
    params = scenes.pbf_params((3.2 * scale, 1.6 * scale, 1.6 * scale), pos.shape[0], h=h, spacing=s,
                               gravity=-9.8 * scale)
    return params, pos.astype(np.float32), np.zeros_like(pos, dtype=np.float32)


def _oracle(params, pos, vel):
    o = O.Oracle(O.params_from_pbf(params), mode=O.JACOBI, threads=THREADS)
    o.set_state(pos, pos, vel)
    return o


def _same(g, o, what):
    st = o.state()
    assert np.array_equal(g.positions(), st["x"].astype(np.float32)), what + ": positions"
    assert np.array_equal(g.velocities(), st["v"].astype(np.float32)), what + ": velocities"


@pytest.mark.parametrize("scale", [1e-3, 2.5e-3, 1.0])
def test_scaled_dam_break_steps_and_solve(pbf, scenes, scale):
    params, pos, vel = _scene(scenes, scale, seed=61)
    static = _static_ranges(params.smoothing_radius)
    print(f"h = {params.smoothing_radius:.3g}: range tests discharged on the host: {static}", flush=True)
    assert static == (scale >= 2.5e-3)
    g = pbf.Particles(params)
    g.add(pos, pos, vel)
    o = _oracle(params, pos, vel)
    for k in (1, 3):
        g.step(k if k == 1 else 2)
        o.step(k if k == 1 else 2)
        _same(g, o, f"scale {scale}, substep {k}")
    used, hist, lsum, kept = o.solve(0.0, 6)
    g.solve(0.0, 6)
    info = g.solve_info()
    assert info["iterations"] == used
    assert np.array_equal(info["max_error_history"], hist), "error history"
    _same(g, o, f"scale {scale}, a 6-slot solve")
    g.close(); o.close()
