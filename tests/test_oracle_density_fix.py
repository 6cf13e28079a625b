"""Closed-form pins of the near-coincidence fix in compute_densities (CPU only).

Simulation/Source/particles.h:300-309: while particle p walks its list, a
neighbour q closer than 1e-8 moves x*_p by -v_p dt / 50, in place, before the
distance that enters the density is taken:

    for q in neighbors[p]:
        if |x*_p - x*_q| < 1e-8:  x*_p -= v_p dt / 50
        rho_p += m W(|x*_p - x*_q|)

What that fixes, for particles placed at one point x with distinct velocities
(delta_p = |v_p| dt / 50, far above 1e-8 here):
  * p moves once per coincident neighbour that is still within 1e-8 of the
    moved x*_p, so with two coincident neighbours it moves once (the second
    test sees the moved x*_p), and a velocity so small that the moved x*_p
    stays within 1e-8 moves it once per neighbour;
  * the distance in the density is the moved one: rho_p = m (W(0) + W(delta_p));
  * on the serial backend the move is written in place, so a later particle q
    reads p's moved x* and no longer sees a coincident neighbour; with Jacobi
    semantics (every pass reads x* as it was when the pass began) q still sees
    x and moves as well.

Each expected value below comes from those statements and the textbook Poly6
kernel (Mueller et al. 2003) in float64.  test_planted_mistakes_fail builds
copies of the oracle with one mistake each (the sign of the move, the unmoved
x*_p in the distance, one move per pass, Jacobi reading moved neighbours, the
serial backend not writing in place, a wrong divisor) and shows that every
copy fails these checks.
"""
import math
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pbf_oracle.cpp")

H = float(np.float32(0.1))
M = 0.125
DT = float(np.float32(1.0 / 120.0))
X = np.array([0.5, 0.5, 0.5])


def _params():
    p = O.OrcParams()
    p.h, p.s = H, 1.2 * H
    p.rho0, p.mass, p.g, p.gamma, p.eps, p.visc, p.dt, p.vort = 1000.0, M, -9.8, 0.0, 600.0, 0.01, DT, 0.0005
    p.max_speed = 0.5 * H * 4 / DT
    for i in range(3):
        p.bmin[i], p.bmax[i], p.mover[i] = 0.0, 1.0, -10.0
    p.iters, p.cap = 4, 60
    return p


def W(r):
    """Poly6 (Mueller et al. 2003 eq. 20) for r <= h."""
    return 315.0 / (64.0 * math.pi * H ** 9) * (H * H - r * r) ** 3


def _density_pass(vel, mode, so=None):
    """Particles all at X with velocities vel: neighbour lists, then one
    compute_densities pass in double precision."""
    n = len(vel)
    pts = np.tile(X, (n, 1))
    o = O.Oracle(_params(), double=True, mode=mode, so=so)
    o.set_state(pts, pts, np.asarray(vel, np.float64))
    o.pass_(O.FIND)
    nb = o.neighbors()
    o.pass_(O.DENSITY)
    st = o.state()
    hacks = o.counters()["hacks"]
    o.close()
    return st, hacks, nb


V = [np.array([1.2, 0.0, 0.0]), np.array([0.0, -0.6, 0.3]), np.array([0.2, 0.4, -0.9])]


def _delta(v):
    return np.linalg.norm(v) * DT / 50.0


def check_pair_jacobi(so=None):
    """Both particles see the other at distance 0 and move by -v dt/50."""
    st, hacks, nb = _density_pass(V[:2], O.JACOBI, so)
    assert list(nb["counts"]) == [1, 1]
    np.testing.assert_allclose(st["xs"], [X - V[0] * DT / 50.0, X - V[1] * DT / 50.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(st["rho"], [M * (W(0) + W(_delta(V[0]))), M * (W(0) + W(_delta(V[1])))],
                               rtol=1e-12)
    assert hacks == 2


def check_pair_serial(so=None):
    """Particle 0 moves in place; particle 1 reads the moved x*_0, is not
    coincident, stays at X, and its density uses |x*_0 - X| = delta_0."""
    st, hacks, _ = _density_pass(V[:2], O.SERIAL, so)
    np.testing.assert_allclose(st["xs"], [X - V[0] * DT / 50.0, X], rtol=0, atol=1e-15)
    np.testing.assert_allclose(st["rho"], [M * (W(0) + W(_delta(V[0]))), M * (W(0) + W(_delta(V[0])))],
                               rtol=1e-12)
    assert hacks == 1


def check_triple_jacobi(so=None):
    """Three coincident particles: each moves once (after the move its second
    coincident neighbour lies delta_p away), and both terms use delta_p."""
    st, hacks, nb = _density_pass(V, O.JACOBI, so)
    assert list(nb["counts"]) == [2, 2, 2]
    np.testing.assert_allclose(st["xs"], [X - v * DT / 50.0 for v in V], rtol=0, atol=1e-15)
    np.testing.assert_allclose(st["rho"], [M * (W(0) + 2.0 * W(_delta(v))) for v in V], rtol=1e-12)
    assert hacks == 3


def check_triple_serial(so=None):
    """Serial order: 0 moves (lists in index order: 0 sees 1, 2); 1 sees the
    moved 0 at delta_0, then 2 coincident and moves; 2 sees both moved."""
    st, hacks, _ = _density_pass(V, O.SERIAL, so)
    d0, d1 = _delta(V[0]), _delta(V[1])
    np.testing.assert_allclose(st["xs"], [X - V[0] * DT / 50.0, X - V[1] * DT / 50.0, X], rtol=0, atol=1e-15)
    np.testing.assert_allclose(st["rho"], [M * (W(0) + 2.0 * W(d0)), M * (W(0) + W(d0) + W(d1)),
                                           M * (W(0) + W(d0) + W(d1))], rtol=1e-12)
    assert hacks == 2


def check_tiny_velocity_moves_once_per_neighbour(so=None):
    """|v| dt / 50 = 5e-11 < 1e-8: after the first move x*_p is still within
    1e-8 of the next coincident neighbour, so it moves again: twice in all."""
    vt = [np.array([3e-7, 0.0, 0.0]), np.array([0.0, -3e-7, 0.0]), np.array([0.0, 0.0, 3e-7])]
    st, hacks, _ = _density_pass(vt, O.JACOBI, so)
    np.testing.assert_allclose(st["xs"], [X - 2.0 * v * DT / 50.0 for v in vt], rtol=0, atol=1e-20)
    assert hacks == 6


CHECKS = [check_pair_jacobi, check_pair_serial, check_triple_jacobi, check_triple_serial,
          check_tiny_velocity_moves_once_per_neighbour]


@pytest.mark.parametrize("check", CHECKS, ids=lambda f: f.__name__)
def test_density_fix_closed_form(check):
    check()


# Planted mistakes: (name, [(text in pbf_oracle.cpp, replacement)]).  Each
# text must occur in compute_densities exactly once (the `rd[...]` reads are
# that pass's own).
MUTANTS = [
    ("plus_sign", [("if (mag(sub(xp, rd[q])) < 0.00000001) {\n          dec(xp,",
                    "if (mag(sub(xp, rd[q])) < 0.00000001) {\n          acc(xp,")]),
    ("unmoved_position_in_r", [("V3<R> xp = rd[p];\n      long long lh = 0;",
                                "V3<R> xp = rd[p];\n      const V3<R> xp0 = xp;\n      long long lh = 0;"),
                               ("const R r = mag(sub(xp, rd[q]));", "const R r = mag(sub(xp0, rd[q]));")]),
    ("one_move_per_pass", [("if (mag(sub(xp, rd[q])) < 0.00000001) {",
                            "if (lh == 0 && mag(sub(xp, rd[q])) < 0.00000001) {")]),
    ("jacobi_reads_moved", [("const std::vector<V3<R>>& rd = (mode == ORC_JACOBI) ? src : xs;",
                             "const std::vector<V3<R>>& rd = xs;")]),
    ("serial_not_in_place", [("          if (mode == ORC_SERIAL) xs[p] = xp;\n", ""),
                             ("        if (mode == ORC_SERIAL) xp = xs[p];\n", "")]),
    ("divide_by_5", [("if (mag(sub(xp, rd[q])) < 0.00000001) {\n          dec(xp, divs(muls(v[p], dt), (R)50.0));",
                      "if (mag(sub(xp, rd[q])) < 0.00000001) {\n          dec(xp, divs(muls(v[p], dt), (R)5.0));")]),
]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(SRC).read()
    d = tmp_path_factory.mktemp("oracle_mutants")
    jobs = []
    for name, edits in MUTANTS:
        s = src
        for old, new in edits:
            assert s.count(old) == 1, f"{name}: the text to replace must occur once in the oracle"
            s = s.replace(old, new)
        path = d / f"{name}.cpp"
        path.write_text(s)
        jobs.append((name, str(path), str(d / f"lib{name}.so")))

    def build(job):
        _, cpp, so = job
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                        "-shared", "-o", so, cpp], check=True)
        return so

    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        libs = list(ex.map(build, jobs))
    return {name: so for (name, _, _), so in zip(jobs, libs)}


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_planted_mistakes_fail(mutant_libs, name):
    failed = []
    for check in CHECKS:
        try:
            check(mutant_libs[name])
        except AssertionError:
            failed.append(check.__name__)
    assert failed, f"the planted mistake {name} passes every closed-form check"
