"""Pins of the CPU oracle (CPU only).

The reference cannot be built here (Common/parameters.h needs Boost, which
the image lacks), so the oracle is pinned by (1) the known answers of the
reference's own tests (tests/golden/known_answers.json), (2) library
routines for the ordering semantics (numpy stable argsort / searchsorted),
(3) brute force on small inputs, and (4) what the mathematics fixes:
textbook closed forms (Mueller et al. 2003 kernels, Macklin & Mueller 2013
PBF equations), invariants (momentum of dp, rigid rotation) and special
cases.  Each test names the mistake it would catch.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def f32(x):
    """The oracle's parameter block is float: closed forms use the float-rounded constant."""
    return float(np.float32(x))


def mk(h=0.1, s=None, box=(1.0, 1.0, 1.0), rho0=1000.0, mass=0.125, g=-9.8, dt=1 / 120, eps=600.0, visc=0.01,
       vort=0.0005, gamma=0.0, iters=4, cap=60, max_speed=None, mover=(-10.0, -10.0, -10.0)):
    p = O.OrcParams()
    p.h = h
    p.s = 1.2 * h if s is None else s
    p.rho0, p.mass, p.g, p.gamma, p.eps, p.visc, p.dt, p.vort = rho0, mass, g, gamma, eps, visc, dt, vort
    p.max_speed = 0.5 * h * iters / dt if max_speed is None else max_speed
    for i in range(3):
        p.bmin[i] = 0.0
        p.bmax[i] = box[i]
        p.mover[i] = mover[i]
    p.iters, p.cap = iters, cap
    return p


def lattice_points(nx, ny, nz, spacing, origin=(0.0, 0.0, 0.0)):
    s = np.float32(spacing)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float32)
    return (idx * s + np.asarray(origin, np.float32) + s / np.float32(2)).astype(np.float32)


# ---------------------------------------------------------------- binning
def test_bin_ids_known_answers():
    """neighbor_test.cpp:94-121.  Catches a wrong linearisation or a missing +s shift."""
    ka = KA["bin_ids_3d"]
    o = O.Oracle(mk(h=1.0, s=ka["bin_spacing"], box=ka["box_max"]))
    assert list(o.dims()) == ka["grid_dims"]
    assert [o.bin_id(p) for p in ka["points"]] == ka["bin_ids"]


def test_neighbor_counts_known_answers():
    """neighbor_test.cpp:211-298.  Catches <= vs <, self not skipped, a missing cell."""
    ka = KA["neighbor_counts_3d"]
    pts = lattice_points(*ka["lattice"], ka["spacing"])
    o = O.Oracle(mk(h=1.0, s=1.0, box=(5.0, 5.0, 5.0)))
    o.set_state(pts)
    o.pass_(O.FIND)
    c = o.neighbors()["counts"].reshape(10, 10, 10)  # [k][j][i]
    inner = slice(1, 9)
    assert np.all(c[inner, inner, inner] == ka["interior"])
    assert np.all(c[0, inner, inner] == ka["face"]) and np.all(c[inner, inner, 9] == ka["face"])
    assert np.all(c[inner, 0, 0] == ka["edge"]) and np.all(c[9, 9, inner] == ka["edge"])
    for k in (0, 9):
        for j in (0, 9):
            for i in (0, 9):
                assert c[k, j, i] == ka["corner"]


def _scene(n_side=9, spacing=0.05, seed=0, box=(1.0, 0.8, 0.8), amp=0.004):
    rng = np.random.default_rng(seed)
    pts = lattice_points(n_side, n_side - 2, n_side - 1, spacing, origin=(0.1, 0.05, 0.05))
    return (pts + rng.uniform(-amp, amp, pts.shape).astype(np.float32)).astype(np.float32)


def _keys_f32(pts, s, dims):
    s = np.float32(s)
    f = np.floor((pts + s) / s).astype(np.int64)
    return (dims[0] * dims[1] * f[:, 2] + (f[:, 1] * dims[0] + f[:, 0])).astype(np.uint64)


def test_sort_and_bounds_match_stable_sort():
    """sort_by_key must be stable (Thrust stable_sort_by_key); bounds = searchsorted."""
    pts = _scene(seed=1)
    o = O.Oracle(mk(box=(1.0, 0.8, 0.8)))
    o.set_state(pts)
    o.pass_(O.FIND)
    nb = o.neighbors()
    keys = _keys_f32(pts, o.params.s, o.dims())
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(nb["ids"], order.astype(np.uint64))
    assert np.array_equal(nb["keys"], keys[order])
    cells = np.arange(len(nb["begin"]))
    assert np.array_equal(nb["begin"], np.searchsorted(keys[order], cells, side="left"))
    assert np.array_equal(nb["end"], np.searchsorted(keys[order], cells, side="right"))


def _walk_rank(pts, s, p, q):
    """Position of q's cell in p's 27-cell walk (x offset outermost, z innermost)."""
    cp = np.floor((pts[p].astype(np.float64) + s) / s)
    cq = np.floor((pts[q].astype(np.float64) + s) / s)
    d = (cq - cp).astype(int) + 1
    return d[0] * 9 + d[1] * 3 + d[2]


def test_neighbor_lists_brute_force_and_order():
    """Accepted set = brute force d^2 < s^2 (float32 arithmetic); order = walk rank then index;
    cap truncation keeps the first cap entries.  Catches a wrong walk order or cap rule."""
    pts = _scene(seed=2)
    n = pts.shape[0]
    o = O.Oracle(mk(box=(1.0, 0.8, 0.8), cap=200))
    o.set_state(pts)
    o.pass_(O.FIND)
    full = o.neighbors()
    s32 = np.float32(o.params.s)
    ss = s32 * s32
    for p in range(0, n, 7):
        d = (pts[p][None, :] - pts).astype(np.float32)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        expect = [q for q in np.nonzero(d2 < ss)[0] if q != p]
        expect.sort(key=lambda q: (_walk_rank(pts, float(s32), p, q), q))
        got = full["lists"][p, : full["counts"][p]].tolist()
        assert got == expect, p
    # cap truncation: the first 20 of the full list, overflow counted
    o2 = O.Oracle(mk(box=(1.0, 0.8, 0.8), cap=20))
    o2.set_state(pts)
    o2.pass_(O.FIND)
    cut = o2.neighbors()
    assert np.array_equal(cut["counts"], np.minimum(full["counts"], 20))
    for p in range(n):
        c = cut["counts"][p]
        assert np.array_equal(cut["lists"][p, :c], full["lists"][p, :c])
    assert o2.counters()["overflow"] == int(np.maximum(full["counts"] - 20, 0).sum())


def _cells_f32(pts, s32):
    """Cell coordinates as the reference computes them in float: floor((x + s) / s) (neighbors.h:108-114)."""
    return np.floor((pts.astype(np.float32) + s32) / s32).astype(np.int64)


def _brute_lists(pts, s, cap):
    """Every q != p with d^2 < s^2 (float32 arithmetic, the reference's order of the sum), ordered by
    walk rank (x offset outermost, z innermost, cells in float32) then index, cut at cap; also the number
    cut."""
    s32 = np.float32(s)
    ss = s32 * s32
    cells = _cells_f32(pts, s32)
    lists, cut = [], 0
    for p in range(pts.shape[0]):
        d = (pts[p][None, :] - pts).astype(np.float32)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        acc = [q for q in np.nonzero(d2 < ss)[0] if q != p]
        rank = lambda q: tuple(int(v) for v in cells[q] - cells[p] + 1)
        assert all(0 <= r <= 2 for q in acc for r in rank(q)), "an accepted neighbour outside the 27 cells"
        acc.sort(key=lambda q: (rank(q), q))
        cut += max(0, len(acc) - cap)
        lists.append(acc[:cap])
    return lists, cut


@pytest.mark.parametrize("case", ["empty", "single", "on_cell_faces", "coincident", "one_cell", "random"])
@pytest.mark.parametrize("cap", [1, 2, 7, 60])
def test_neighbor_lists_edge_cases_brute_force(case, cap):
    """fill_neighbors (neighbors.h:197-230) against brute force on the inputs where a walk goes wrong
    first: no particles, one, particles exactly on cell faces (the floor of (x + s)/s), coincident
    duplicates (d^2 = 0 is accepted, self is not), everything in one cell, and a random cloud; with caps
    that cut most lists.  The overflow counter equals the entries cut."""
    rng = np.random.default_rng([cap, sum(map(ord, case))])
    s = 0.125  # exact in binary: points on the cell faces sit exactly s apart (d^2 == s^2 is rejected)
    if case == "empty":
        pts = np.zeros((0, 3), np.float32)
    elif case == "single":
        pts = np.array([[0.5, 0.5, 0.5]], np.float32)
    elif case == "on_cell_faces":
        g = (np.arange(2, 7) * np.float32(s)).astype(np.float32)
        pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    elif case == "coincident":
        base = rng.uniform(0.3, 0.6, size=(20, 3)).astype(np.float32)
        pts = np.concatenate([base, base, base[:7]]).astype(np.float32)
    elif case == "one_cell":
        pts = (0.25 + rng.uniform(0.0, 0.1, size=(90, 3))).astype(np.float32)
    else:
        pts = rng.uniform(0.05, 0.95, size=(400, 3)).astype(np.float32)
    o = O.Oracle(mk(box=(1.0, 1.0, 1.0), s=s, cap=cap))
    o.set_state(pts)
    o.pass_(O.FIND)
    nb = o.neighbors()
    expect, cut = _brute_lists(pts, o.params.s, cap)
    assert nb["counts"].tolist() == [len(x) for x in expect]
    for p in range(pts.shape[0]):
        assert nb["lists"][p, : nb["counts"][p]].tolist() == expect[p], p
    assert o.counters()["overflow"] == cut


def test_stop_at_cap_keeps_lists_and_steps():
    """The checker shortcut that ends a walk once its list is full (for the full-size GPU tests'
    collapsed states) keeps every list, count and step result; only the overflow count is lost."""
    pts = _scene(seed=2)
    res = []
    for stop in (False, True):
        o = O.Oracle(mk(box=(1.0, 0.8, 0.8), cap=20))
        o.set_stop_at_cap(stop)
        o.set_state(pts)
        o.pass_(O.FIND)
        res.append((o.neighbors(), o.counters()["overflow"]))
        o.step(3)
        res[-1] += (o.state(),)
    assert res[0][1] > 0 and res[1][1] == 0
    for k in ("counts", "lists", "ids", "begin", "end"):
        assert np.array_equal(res[0][0][k], res[1][0][k]), k
    for k in ("x", "v", "rho", "lam"):
        assert np.array_equal(res[0][2][k], res[1][2][k]), k


# ---------------------------------------------------------------- kernels
def test_kernel_zeros_known_answers():
    """kernels_test.cpp: Poly6 and C_Spline are 0 at r = h and beyond; Del_Spikey 0 beyond h."""
    ka = KA["kernels"]
    h = ka["h"]
    for dbl in (False, True):
        o = O.Oracle(mk(h=h), double=dbl)
        for r in ka["poly6_zero_at"]:
            assert o.poly6(r) == pytest.approx(0.0, abs=1e-6)
        for r in ka["cspline_zero_at"]:
            assert o.cspline(r) == pytest.approx(0.0, abs=1e-6)
        q = np.full(3, h / math.sqrt(3.0) + 0.001)
        assert np.linalg.norm(o.spiky(np.zeros(3), q)) == 0.0


def test_poly6_normalisation():
    """kernels_test.cpp:86-102: the integral of Poly6 over a 101^3 grid is 1 within 0.1%."""
    ka = KA["kernels"]
    h = ka["h"]
    dim = ka["poly6_integral_grid"]
    ds = 2.0 * h / (dim - 1)
    ax = (np.arange(dim) - dim // 2) * ds
    r2 = (ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2).ravel()
    uniq, cnt = np.unique(np.round(r2 / ds ** 2).astype(np.int64), return_counts=True)
    for dbl in (False, True):
        o = O.Oracle(mk(h=h), double=dbl)
        w = np.array([o.poly6(math.sqrt(u) * ds) for u in uniq])
        total = float((w * cnt).sum() * ds ** 3)
        assert total == pytest.approx(ka["poly6_integral"], rel=ka["poly6_integral_rel_tol"])


def test_poly6_textbook_closed_form():
    """Mueller et al. 2003: W = 315/(64 pi h^9) (h^2 - r^2)^3.  Catches a wrong norm or power."""
    h = f32(0.1)
    o = O.Oracle(mk(h=h), double=True)
    for r in (0.0, 0.013, 0.05, 0.0999):
        assert o.poly6(r) == pytest.approx(315.0 / (64 * math.pi * h ** 9) * (h * h - r * r) ** 3, rel=1e-12)


def test_cohesion_spline_closed_form_properties():
    """Akinci et al. 2013 eq. 2 (cohesion spline): C = 32/(pi h^9) (h - r)^3 r^3 for h/2 < r <= h and
    32/(pi h^9) (2 (h - r)^3 r^3 - h^6/64) for r <= h/2.  What the definition fixes:
    C(h/2) = 1/(2 pi h^3) = -C(0) (a wrong norm power or a dropped or sign-flipped h^6/64 term fails);
    the two branches meet at h/2 (a wrong factor 2 or swapped branches fail); the one root below h/2 is
    at r/h = (1 - sqrt(1 - 2^(-1/3)))/2, where (h - r) r = 2^(-7/3) h^2; C rises on [0, h/2] and
    falls on [h/2, h]."""
    for h in (f32(0.1), f32(0.05)):
        o = O.Oracle(mk(h=h), double=True)
        top = 1.0 / (2.0 * math.pi * h ** 3)
        assert o.cspline(0.5 * h) == pytest.approx(top, rel=1e-12)
        assert o.cspline(0.0) == pytest.approx(-top, rel=1e-12)
        assert o.cspline(0.5 * h * (1 + 1e-9)) == pytest.approx(top, rel=1e-7)
        root = 0.5 * (1.0 - math.sqrt(1.0 - 2.0 ** (-1.0 / 3.0))) * h
        assert o.cspline(root * (1 - 1e-6)) < 0.0 < o.cspline(root * (1 + 1e-6))
        assert abs(o.cspline(root)) <= 1e-9 * top
        rs = np.linspace(0.0, h, 401)
        c = np.array([o.cspline(r) for r in rs])
        assert np.all(np.diff(c[:201]) > 0) and np.all(np.diff(c[200:]) < 0)
        assert c[-1] == pytest.approx(0.0, abs=1e-12 * top)


def test_spiky_gradient_is_gradient_of_spiky_kernel():
    """Del_Spikey(p, q) = grad_p W_spiky(|p - q|), W_spiky = 15/(pi h^6) (h - r)^3
    (Mueller et al. 2003); central differences in double.  Catches sign and norm errors."""
    h = f32(0.1)
    o = O.Oracle(mk(h=h), double=True)
    W = lambda r: 15.0 / (math.pi * h ** 6) * (h - r) ** 3 if r <= h else 0.0
    rng = np.random.default_rng(3)
    for _ in range(20):
        p = rng.uniform(0.2, 0.3, 3)
        q = p + rng.normal(size=3) * 0.03
        r = np.linalg.norm(p - q)
        if r > 0.95 * h:
            continue
        num = np.zeros(3)
        e = 1e-7
        for i in range(3):
            dp = np.zeros(3)
            dp[i] = e
            num[i] = (W(np.linalg.norm(p + dp - q)) - W(np.linalg.norm(p - dp - q))) / (2 * e)
        np.testing.assert_allclose(o.spiky(p, q), num, rtol=1e-5, atol=1e-6 * np.abs(num).max())


# ---------------------------------------------------------------- particle passes
def test_external_forces_and_predict_known_answers():
    """particle_test.cpp:122-225: v.y += g dt; x* = x + v dt; out-of-bounds clamped to the minimum."""
    ka = KA["particles"]
    pts = lattice_points(4, 4, 4, 0.5, origin=(2.0, 2.0, 2.0))
    o = O.Oracle(mk(h=0.1, box=ka["box_max"], dt=ka["time_step"], g=ka["gravity"]))
    o.set_state(pts)
    o.pass_(O.FORCES)
    v = o.state()["v"]
    assert np.allclose(v[:, 1], ka["gravity"] * ka["time_step"]) and np.all(v[:, [0, 2]] == 0)
    o2 = O.Oracle(mk(h=0.1, box=ka["box_max"], dt=ka["time_step"]))
    vel = np.tile(np.array(ka["predict_velocity"], np.float32), (pts.shape[0], 1))
    o2.set_state(pts, pts, vel)
    o2.pass_(O.PREDICT)
    np.testing.assert_allclose(o2.state()["xs"], pts + vel * ka["time_step"], rtol=1e-6)
    x0 = pts.copy()
    x0[0] = 0.0
    v0 = vel.copy()
    v0[0] = [-1.0, -2.0, -3.0]
    o2.set_state(x0, x0, v0)
    o2.pass_(O.PREDICT, 0, 1)
    assert np.all(o2.state()["xs"][0] == 0.0)


def test_mover_pushes_particle_to_sphere_surface():
    """apply_boundary_conditions (particles.h:514-522): a point inside the r = 0.2 mover ends on its surface."""
    o = O.Oracle(mk(box=(2.0, 2.0, 2.0), mover=(1.0, 1.0, 1.0)), double=True)
    p = o.boundary([1.05, 1.02, 0.97])
    assert np.linalg.norm(p - 1.0) == pytest.approx(0.2, rel=1e-6)
    d_in = np.array([0.05, 0.02, -0.03])
    assert np.allclose((p - 1.0) / 0.2, d_in / np.linalg.norm(d_in))


def test_parameters_derivation_known_answers(pbf):
    """parameters_test.cpp:28-84 through pbf_params_derive (host logic of the library)."""
    ka = KA["parameters"]
    p = pbf.Params()
    p.particle_rest_spacing = (ka["initial_fluid_volume"] / ka["global_particle_count"]) ** (1.0 / 3.0)
    p.solve_step_count = ka["solve_step_count"]
    p.time_step = ka["time_step"]
    pbf.derive(p)
    assert p.particle_rest_spacing == pytest.approx(ka["particle_rest_spacing"], rel=1e-4)
    assert p.smoothing_radius == pytest.approx(ka["smoothing_radius_over_spacing"] * p.particle_rest_spacing)
    assert p.neighbor_bin_spacing == pytest.approx(ka["bin_spacing_over_h"] * p.smoothing_radius)
    assert p.rest_density == pytest.approx(ka["rest_density"], rel=1e-6)
    assert p.rest_mass == ka["rest_mass"]
    assert p.max_speed == pytest.approx(ka["max_speed"], rel=1e-5)
    assert p.max_neighbors == 60


def test_density_mean_within_five_percent_known_answer():
    """particle_test.cpp:245-265: 30^3 lattice (particle_test.ini), mean density within 5% of rest."""
    ka = KA["particles"]
    s = (8.0 / ka["global_particle_count"]) ** (1.0 / 3.0)
    s32 = float(np.float32(s))
    n_side = int(math.floor(np.float32(2.0) / np.float32(s32)))
    assert n_side ** 3 == ka["constructed_3d"]
    pts = lattice_points(n_side, n_side, n_side, s32, origin=ka["fluid_min"])
    h = 1.8 * s32
    o = O.Oracle(mk(h=h, s=1.2 * h, box=ka["box_max"], rho0=1.0 / s32 ** 3, mass=1.0), threads=4)
    o.set_state(pts)
    o.pass_(O.FIND)
    o.pass_(O.DENSITY)
    rho = o.state()["rho"]
    assert abs(rho.mean() - 1.0 / s32 ** 3) < ka["mean_density_rel_tol"] / s32 ** 3


def test_density_two_particle_closed_form():
    """rho_p = m (W(0) + W(r)) for an isolated pair; catches a missing self term."""
    h = f32(0.1)
    m = 0.125
    o = O.Oracle(mk(h=h, box=(1, 1, 1), mass=m), double=True)
    pts = np.array([[0.5, 0.5, 0.5], [0.53, 0.51, 0.49]])
    o.set_state(pts)
    o.pass_(O.FIND)
    o.pass_(O.DENSITY)
    r = np.linalg.norm(pts[0] - pts[1])
    W = lambda r: 315.0 / (64 * math.pi * h ** 9) * (h * h - r * r) ** 3
    np.testing.assert_allclose(o.state()["rho"], m * (W(0) + W(r)), rtol=1e-12)


def _textbook_pbf(pts, lists, counts, h, rho0, mass, eps):
    """Macklin & Mueller 2013 eq. 2, 8-12 in float64, written per constraint
    gradient w.r.t. each particle k (not the reference's accumulation form).
    The constraint is clamped at 0 as the reference does (particles.h:325)."""
    n = pts.shape[0]
    W = lambda r: np.where(r <= h, 315.0 / (64 * math.pi * h ** 9) * (h * h - r * r) ** 3, 0.0)

    def gradW(d):  # gradient of the Spiky kernel w.r.t. the first point, d = p_i - p_j
        r = np.linalg.norm(d, axis=-1, keepdims=True)
        g = -45.0 / (math.pi * h ** 6) * (h - r) ** 2 * d / (r + np.finfo(float).eps)
        return np.where(r <= h, g, 0.0)

    rho = np.empty(n)
    lam = np.empty(n)
    for i in range(n):
        nb = lists[i, : counts[i]].astype(int)
        d = pts[i] - pts[nb]
        rho[i] = mass * (W(0.0) + W(np.linalg.norm(d, axis=1)).sum())
        C = max(rho[i] / rho0 - 1.0, 0.0)
        grads_k = -gradW(d) / rho0          # k = j
        grad_i = gradW(d).sum(0) / rho0     # k = i
        denom = (grads_k ** 2).sum() + (grad_i ** 2).sum() + eps
        lam[i] = -C / denom
    dp = np.zeros((n, 3))
    for i in range(n):
        nb = lists[i, : counts[i]].astype(int)
        dp[i] = ((lam[i] + lam[nb])[:, None] * gradW(pts[i] - pts[nb])).sum(0) / rho0
    return rho, lam, dp


def _compressed_cluster(seed=4):
    rng = np.random.default_rng(seed)
    pts = lattice_points(6, 6, 6, 0.04, origin=(0.4, 0.4, 0.4)).astype(np.float64)  # 20% compressed
    return pts + rng.uniform(-0.004, 0.004, pts.shape)


def test_lambda_and_dp_match_textbook_pbf():
    """Catches: dropped k = i term, wrong sign, missing 1/rho0, lambda_p vs lambda_p + lambda_q."""
    pts = _compressed_cluster()
    h, rho0, m, eps = f32(0.1), 1000.0, 0.125, 600.0
    o = O.Oracle(mk(h=h, box=(1.0, 1.0, 1.0), rho0=rho0, mass=m, eps=eps, cap=400), double=True)
    o.set_state(pts)
    o.pass_(O.FIND)
    o.pass_(O.DENSITY)
    o.pass_(O.LAMBDA)
    o.pass_(O.DP)
    st = o.state()
    nb = o.neighbors()
    rho, lam, dp = _textbook_pbf(pts, nb["lists"], nb["counts"], h, rho0, m, eps)
    assert (lam < 0).sum() > 100, "cluster must be compressed"
    np.testing.assert_allclose(st["rho"], rho, rtol=1e-12)
    np.testing.assert_allclose(st["lam"], lam, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(st["scratch"], dp, rtol=1e-9, atol=1e-12)


def test_dp_conserves_momentum_and_separates_compressed_pair():
    """Sum of dp over symmetric lists is 0 ((lambda_p + lambda_q) symmetric, gradient antisymmetric);
    a compressed pair is pushed apart."""
    pts = _compressed_cluster(seed=5)
    o = O.Oracle(mk(h=0.1, box=(1.0, 1.0, 1.0), cap=400), double=True)
    o.set_state(pts)
    for pid in (O.FIND, O.DENSITY, O.LAMBDA, O.DP):
        o.pass_(pid)
    dp = o.state()["scratch"]
    assert np.abs(dp.sum(0)).max() <= 1e-12 * np.abs(dp).sum()
    pair = np.array([[0.5, 0.5, 0.5], [0.52, 0.5, 0.5]])
    o2 = O.Oracle(mk(h=0.1, box=(1, 1, 1), rho0=10.0, mass=0.125), double=True)  # rho > rho0
    o2.set_state(pair)
    for pid in (O.FIND, O.DENSITY, O.LAMBDA, O.DP):
        o2.pass_(pid)
    d = o2.state()["scratch"]
    assert d[0, 0] < 0 < d[1, 0] and d[0, 0] == pytest.approx(-d[1, 0])


def test_update_velocities_threshold():
    """v = (x* - x)/dt, zeroed when |v|^2 < 1e-6 max_speed (compared in double)."""
    dt = f32(0.01)
    o = O.Oracle(mk(dt=dt, max_speed=2.0), double=True)
    x = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    thr = math.sqrt(1e-6 * 2.0)
    xs = x + np.array([[1.01 * thr * dt, 0, 0], [0.99 * thr * dt, 0, 0]])
    o.set_state(x, xs)
    o.pass_(O.UPDATE_V)
    v = o.state()["v"]
    assert v[0, 0] == pytest.approx(1.01 * thr) and np.all(v[1] == 0.0)


def _pair_with_density(v0, v1, rho_val=900.0):
    pts = np.array([[0.5, 0.5, 0.5], [0.53, 0.5, 0.5]])
    return pts, np.array([v0, v1], dtype=np.float64), np.full(2, rho_val)


def test_viscosity_closed_form_jacobi_and_serial():
    """XSPH v_p += c (v_q - v_p) W / rho_q.  Serial mode updates in index order
    (particles.h:439-446): particle 1 sees particle 0's new velocity."""
    h, c = f32(0.1), f32(0.01)
    pts, vel, rho = _pair_with_density([1.0, 0.0, 0.0], [0.0, 2.0, 0.0])
    W = 315.0 / (64 * math.pi * h ** 9) * (h * h - 0.03 ** 2) ** 3
    for mode in (O.JACOBI, O.SERIAL):
        o = O.Oracle(mk(h=h, visc=c, box=(1, 1, 1)), double=True, mode=mode)
        o.set_state(pts, pts, vel)
        o.set_aux(rho=rho)
        o.pass_(O.FIND)
        o.pass_(O.VISC)
        v = o.state()["v"]
        v0 = vel[0] + c * (vel[1] - vel[0]) * W / rho[1]
        src0 = vel[0] if mode == O.JACOBI else v0
        v1 = vel[1] + c * (src0 - vel[1]) * W / rho[0]
        np.testing.assert_allclose(v, [v0, v1], rtol=1e-12)
    # a uniform field is a fixed point
    o = O.Oracle(mk(h=h, box=(1, 1, 1)), double=True)
    pts = _compressed_cluster()
    o.set_state(pts, pts, np.tile([0.3, -0.2, 0.1], (pts.shape[0], 1)))
    o.set_aux(rho=np.full(pts.shape[0], 1000.0))
    o.pass_(O.FIND)
    o.pass_(O.VISC)
    np.testing.assert_allclose(o.state()["v"], np.tile([0.3, -0.2, 0.1], (pts.shape[0], 1)), rtol=1e-14)


def test_vorticity_of_rigid_rotation():
    """For v = Omega x r on a symmetric neighbourhood, sum_q (v_q - v_p) x gradW(p - q)
    = -(2/3) sum_q k_q |d_q|^2 Omega, with gradW(p - q) = k_q d_q, d_q = q - p, k_q > 0.
    Catches the cross-product operand order and the gradient sign."""
    h = f32(0.1)
    s = 0.05
    pts = lattice_points(7, 7, 7, s, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    omega = np.array([0.3, -0.7, 1.1])
    vel = np.cross(omega[None, :], pts)
    o = O.Oracle(mk(h=h, box=(1, 1, 1), cap=200), double=True)
    o.set_state(pts, pts, vel)
    o.pass_(O.FIND)
    o.pass_(O.VORT)
    centre = 3 * 49 + 3 * 7 + 3
    d = pts - pts[centre]
    r = np.linalg.norm(d, axis=1)
    inside = (r > 0) & (r <= h)
    k = 45.0 / (math.pi * h ** 6) * (h - r[inside]) ** 2 / (r[inside] + np.finfo(float).eps)
    expect = -(2.0 / 3.0) * (k * r[inside] ** 2).sum() * omega
    np.testing.assert_allclose(o.state()["scratch"][centre], expect, rtol=1e-9)


def test_apply_vorticity_direction():
    """N = eta / (|eta| + eps) has |N| <= 1, so |dv| <= |omega_p| eps dt; a gradient of
    |omega| along +x gives N = +x and v += (N x omega_p) eps dt (catches the cross order
    and the sign of eta)."""
    h, eps_v, dt = f32(0.1), f32(0.0005), f32(1 / 120)
    pts = lattice_points(7, 7, 7, 0.05, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    n = pts.shape[0]
    centre = 3 * 49 + 3 * 7 + 3
    o = O.Oracle(mk(h=h, box=(1, 1, 1), cap=200, vort=eps_v, dt=dt), double=True)
    w = np.tile([0.0, 0.0, 2.0], (n, 1))
    o.set_state(pts, pts, np.zeros((n, 3)))
    o.set_aux(scratch=w)
    o.pass_(O.FIND)
    o.pass_(O.APPLY_VORT)
    assert np.linalg.norm(o.state()["v"], axis=1).max() <= 2.0 * eps_v * dt * (1 + 1e-12)
    w2 = np.zeros((n, 3))
    w2[:, 2] = 1.0 + pts[:, 0]
    o.set_state(pts, pts, np.zeros((n, 3)))
    o.set_aux(scratch=w2)
    o.pass_(O.FIND)
    o.pass_(O.APPLY_VORT)
    vc = o.state()["v"][centre]
    expect = np.cross([1.0, 0.0, 0.0], w2[centre]) * eps_v * dt
    np.testing.assert_allclose(vc, expect, rtol=1e-6, atol=1e-4 * np.linalg.norm(expect))


def test_warm_features_on_a_lattice_closed_forms():
    """Features 4-7 of the learned warm start on a cubic lattice (spacing s, h = 2.2 s) in an affine
    velocity field v = A x.  With gradW(x_p - x_q) = -(45/(pi h^6)) (h - r)^2 (x_p - x_q)/r and
    v_q - v_p = A d (d = x_q - x_p), cubic symmetry (sum f(r) d d^T = (1/3) sum f(r) r^2 I) gives
    feature 7 = m dt sum (v_q - v_p) . grad_q C = -(45/(pi h^6 rho0)) m dt (tr A / 3) sum (h - r)^2 r
    for an interior particle (an expanding field lowers the density; the rotation part of A adds
    nothing).  Feature 6 = mean (y_q - y_p)/h is 0 inside and, on the top layer, s/h times the mean j
    of the integer offsets (i, j, k), j <= 0, i^2 + j^2 + k^2 <= 4; feature 4 = count/32 (32 inside).
    Catches a reversed velocity difference, a dropped rho0, m or dt, the sign of the offset and a
    count over the whole list.  The solve predicts x* = x + v dt, so particles start at lattice - v dt."""
    s, h, rho0, m, dt = 0.05, f32(0.11), 1000.0, f32(0.125), f32(1 / 120)
    nside = 9
    lat = lattice_points(nside, nside, nside, s, origin=(0.3, 0.3, 0.3)).astype(np.float64)
    A = np.array([[0.9, 0.4, -0.2], [-0.4, 1.3, 0.5], [0.2, -0.5, 0.8]])
    centre = np.array([0.5, 0.5, 0.5])
    vel = (lat - centre) @ A.T
    x0 = lat - vel * dt
    o = O.Oracle(mk(h=h, g=0.0, dt=dt, rho0=rho0, mass=m, box=(1, 1, 1), cap=200), double=True)
    o.set_state(x0, x0, vel)
    n = lat.shape[0]
    feat = o.solve(0.0, 1, warm=np.zeros(n), features=True)[4]

    off = np.array([(i, j, k) for i in range(-2, 3) for j in range(-2, 3) for k in range(-2, 3)
                    if 0 < i * i + j * j + k * k <= 4], dtype=np.float64)
    assert off.shape[0] == 32
    r = np.linalg.norm(off, axis=1) * s
    inner = (4 * nside + 4) * nside + 4          # lattice index (4, 4, 4)
    top = (4 * nside + (nside - 1)) * nside + 4  # (4, 8, 4)
    f7 = -45.0 / (math.pi * h ** 6 * rho0) * m * dt * (np.trace(A) / 3.0) * ((h - r) ** 2 * r).sum()
    assert feat[inner, 7] < 0
    assert feat[inner, 7] == pytest.approx(f7, rel=1e-5)
    assert feat[inner, 4] == 1.0 and abs(feat[inner, 6]) <= 1e-6
    below = off[off[:, 1] <= 0]
    assert feat[top, 4] == pytest.approx(below.shape[0] / 32.0, rel=1e-7)
    assert feat[top, 6] == pytest.approx(s * below[:, 1].mean() / h, rel=1e-5)
    assert feat[inner, 5] == pytest.approx(vel[inner, 1] * dt / h, rel=1e-5, abs=1e-7)


def test_isolated_particles_follow_ballistic_trajectory():
    """Whole step with no neighbours: symplectic Euler free fall, y_n = y0 + g dt^2 n(n+1)/2."""
    dt, g = f32(1 / 120), f32(-9.8)
    pts = np.array([[0.2, 0.9, 0.2], [0.7, 0.95, 0.7], [0.2, 0.97, 0.7]])
    o = O.Oracle(mk(dt=dt, g=g, box=(1, 1, 1)), double=True)
    o.set_state(pts)
    n = 10
    o.step(n)
    st = o.state()
    np.testing.assert_allclose(st["x"][:, 1], pts[:, 1] + g * dt * dt * n * (n + 1) / 2, rtol=1e-12)
    np.testing.assert_allclose(st["v"][:, 1], g * dt * n, rtol=1e-9)


def _tension_pair(o, d, rho, vel):
    pts = np.array([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5]])
    o.set_state(pts, pts, vel)
    o.set_aux(rho=rho)
    o.pass_(O.FIND)
    o.pass_(O.TENSION)
    return pts, o.state()["v"]


def test_surface_tension_pair_closed_form():
    """Akinci et al. 2013 eq. 1-6 for two particles: normals n_i = h gradW(x_i - x_j) / rho_j, force
    on i K_ij (-gamma C(r) (x_i - x_j)/r - gamma (n_i - n_j)) with K_ij = 2 rho0 / (rho_i + rho_j), and
    dv_i = F_i / rho_i dt (particles.h:400-431).  The three separations put r in the repulsive and
    attractive parts of the inner spline branch and in the outer branch; the densities differ so that a
    swapped rho_i / rho_j fails.  The spline is written out here from the paper, the spiky gradient is the
    one pinned above."""
    h, gamma, rho0, dt = f32(0.1), f32(0.7), 1000.0, f32(1 / 120)
    rho = np.array([950.0, 1040.0])
    vel = np.array([[0.1, 0.0, 0.0], [0.0, -0.2, 0.0]])

    def spline(r):
        c = 32.0 / (math.pi * h ** 9)
        if r <= 0.5 * h:
            return c * (2.0 * (h - r) ** 3 * r ** 3 - h ** 6 / 64.0)
        return c * (h - r) ** 3 * r ** 3

    for d in (0.02, 0.04, 0.07):
        o = O.Oracle(mk(h=h, gamma=gamma, rho0=rho0, dt=dt, box=(1, 1, 1)), double=True)
        pts, v = _tension_pair(o, d, rho, vel)
        n = [h * o.spiky(pts[0], pts[1]) / rho[1], h * o.spiky(pts[1], pts[0]) / rho[0]]
        K = 2.0 * rho0 / (rho[0] + rho[1])
        for i, j in ((0, 1), (1, 0)):
            rij = pts[i] - pts[j]
            r = float(np.linalg.norm(rij))
            F = K * (-gamma * spline(r) * rij / r - gamma * (n[i] - n[j]))
            np.testing.assert_allclose(v[i], vel[i] + F / rho[i] * dt, rtol=1e-12, atol=1e-15)
        # the cohesion part alone pulls particle 0 towards particle 1 (+x) above the spline's root
        # (0.273 h) and pushes it away below
        coh0 = -gamma * spline(d) * (pts[0] - pts[1]) / d
        assert (coh0[0] > 0) == (d > 0.273 * h)

    # three particles with three densities: a pair cannot tell n_i = h sum gradW / rho_j from a normal
    # divided by rho_i (n_i - n_j is the same), a bent chain can
    pts = np.array([[0.5, 0.5, 0.5], [0.55, 0.52, 0.5], [0.58, 0.47, 0.53]])
    rho3 = np.array([950.0, 1040.0, 1010.0])
    vel3 = np.zeros((3, 3))
    o = O.Oracle(mk(h=h, gamma=gamma, rho0=rho0, dt=dt, box=(1, 1, 1)), double=True)
    o.set_state(pts, pts, vel3)
    o.set_aux(rho=rho3)
    o.pass_(O.FIND)
    o.pass_(O.TENSION)
    v = o.state()["v"]
    dist = np.linalg.norm(pts[:, None] - pts[None], axis=2)
    nb = [[j for j in range(3) if j != i and dist[i, j] <= h] for i in range(3)]
    assert all(len(x) == 2 for x in nb)
    n = [h * sum(o.spiky(pts[i], pts[j]) / rho3[j] for j in nb[i]) for i in range(3)]
    for i in range(3):
        F = np.zeros(3)
        for j in nb[i]:
            rij = pts[i] - pts[j]
            K = 2.0 * rho0 / (rho3[i] + rho3[j])
            F += K * (-gamma * spline(dist[i, j]) * rij / dist[i, j] - gamma * (n[i] - n[j]))
        np.testing.assert_allclose(v[i], F / rho3[i] * dt, rtol=1e-12, atol=1e-15)


def test_surface_tension_conserves_momentum():
    """Each pair's force is antisymmetric (K_ij symmetric, n_i - n_j and the cohesion term odd), so over
    symmetric lists sum_i rho_i dv_i = 0 with any densities.  Catches n_i + n_j, K built from one
    particle's density, and dividing the force by rho_j."""
    pts = _compressed_cluster(seed=7)
    n = pts.shape[0]
    rng = np.random.default_rng(3)
    rho = rng.uniform(900.0, 1100.0, n)
    o = O.Oracle(mk(h=f32(0.1), gamma=f32(2.0), box=(1, 1, 1), cap=400), double=True)
    o.set_state(pts, pts, np.zeros((n, 3)))
    o.set_aux(rho=rho)
    o.pass_(O.FIND)
    assert int(o.counters()["overflow"]) == 0
    o.pass_(O.TENSION)
    mom = rho[:, None] * o.state()["v"]
    assert np.abs(mom).max() > 0
    assert np.abs(mom.sum(0)).max() <= 1e-12 * np.abs(mom).sum()


def test_surface_tension_gamma_zero_is_noop_and_float_double_spread():
    """gamma = 0 leaves velocities unchanged; float and double runs of a small scene stay close."""
    pts = _compressed_cluster()
    o = O.Oracle(mk(box=(1, 1, 1)), double=True)
    vel = np.random.default_rng(0).normal(size=pts.shape)
    o.set_state(pts, pts, vel)
    o.set_aux(rho=np.full(pts.shape[0], 1000.0))
    o.pass_(O.FIND)
    o.pass_(O.TENSION)
    assert np.array_equal(o.state()["v"], vel)
    f = O.Oracle(mk(box=(1, 1, 1)))
    d = O.Oracle(mk(box=(1, 1, 1)), double=True)
    p32 = pts.astype(np.float32)
    f.set_state(p32)
    d.set_state(p32.astype(np.float64))
    f.step(5)
    d.step(5)
    assert np.abs(f.state()["x"] - d.state()["x"]).max() < 1e-4
