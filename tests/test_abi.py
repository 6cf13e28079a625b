"""CPU checks of the C ABI: the library loads, exports every declared entry
point, and its host-side logic (parameter derivation, argument checks) works
without a GPU.  No compute call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pbf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_surface():
    names = declared_functions()
    # one entry per for_each_index body of particles.h / neighbors.h
    for n in ("pbf_apply_external_forces", "pbf_predict_positions", "pbf_find_neighbors", "pbf_compute_densities",
              "pbf_compute_pressure_lambdas", "pbf_compute_pressure_dps", "pbf_update_position_stars",
              "pbf_update_velocities", "pbf_apply_surface_tension", "pbf_apply_viscosity", "pbf_compute_vorticity",
              "pbf_apply_vorticity", "pbf_update_positions", "pbf_sort_by_key_u32", "pbf_lower_bound_counting",
              "pbf_upper_bound_counting", "pbf_step"):
        assert n in names, n


def test_library_exports_every_declared_symbol(pbf):
    L = pbf.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(pbf, monkeypatch, tmp_path):
    """With no built library the package raises instead of falling back to anything else."""
    monkeypatch.setattr(pbf, "_lib", None)
    monkeypatch.setattr(pbf, "LIB_PATH", str(tmp_path / "libpbf.so"))
    with pytest.raises(RuntimeError, match="missing"):
        pbf.lib()
    with pytest.raises(RuntimeError, match="missing"):
        pbf.derive(pbf.Params())


def test_params_derive_defaults(pbf):
    p = pbf.Params()
    p.particle_rest_spacing = 0.05
    p.solve_step_count = 4
    p.time_step = 1 / 120
    pbf.derive(p)
    # parameters.h:127-145
    assert p.particle_radius == pytest.approx(0.025)
    assert p.smoothing_radius == pytest.approx(0.09)
    assert p.neighbor_bin_spacing == pytest.approx(0.108)
    assert p.rest_mass == 1.0
    assert p.rest_density == pytest.approx(8000.0, rel=1e-6)
    assert p.max_speed == pytest.approx(0.5 * 0.09 * 4 * 120, rel=1e-6)
    assert p.max_neighbors == 60


def test_derive_rejects_bad_spacing(pbf):
    p = pbf.Params()
    p.particle_rest_spacing = 0.0
    with pytest.raises(pbf.PBFError) as e:
        pbf.derive(p)
    assert e.value.code == pbf.PBF_ERR_ARG
    assert b"spacing" in pbf.lib().pbf_last_error()


def test_create_rejects_bad_arguments_before_touching_the_device(pbf):
    L = pbf.lib()
    ctx = C.c_void_p()
    p = pbf.Params()
    p.max_particles_local = 0
    assert L.pbf_create(C.byref(p), C.byref(ctx)) == pbf.PBF_ERR_ARG
    p.max_particles_local = 1 << 28  # slots are addressed with 32-bit byte offsets
    assert L.pbf_create(C.byref(p), C.byref(ctx)) == pbf.PBF_ERR_ARG
    assert b"2^28" in pbf.lib().pbf_last_error()
    p.max_particles_local = 10
    p.smoothing_radius = 0.1
    p.neighbor_bin_spacing = 0.12
    p.max_neighbors = 60
    p.time_step = 0.01
    p.boundary_min[0] = 1.0  # the reference bins from the origin (neighbors.h:110-113)
    p.boundary_max[0] = p.boundary_max[1] = p.boundary_max[2] = 2.0
    assert L.pbf_create(C.byref(p), C.byref(ctx)) == pbf.PBF_ERR_UNSUPPORTED
    assert not ctx.value


def test_scene_builders(scenes):
    params, pos, vel = scenes.config_c1(seed=0)
    assert pos.shape == (32768, 3) and pos.dtype == np.float32
    assert params.rest_density == 1000.0 and params.rest_mass == np.float32(0.125)
    assert params.neighbor_bin_spacing == pytest.approx(0.12)
    assert np.all(pos > 0) and np.all(pos[:, 0] < 1.6) and np.all(pos[:, 1] < 1.6)
    # lattice at cell centres plus at most 0.01 h of jitter (particles.h:237-248)
    lat = scenes.lattice((32, 32, 32))
    assert np.abs(pos - lat).max() <= 0.001 + 1e-7
    p2, pos2, _ = scenes.config_c2(seed=0)
    assert pos2.shape[0] == 1_048_576
    assert tuple(p2.boundary_max) == pytest.approx((12.8, 6.4, 6.4))


def test_double_dam_break_and_batch_builders(scenes):
    """The 8M double dam break: two 128 x 256 x 128 blocks on opposite walls of 25.6 x 12.8 x 12.8, cap 64.
    The batch: 256 scenes, counts log-uniform in 4,096..65,536 (up to the lattice rounding), block
    aspect ratios in [0.5, 2], each box twice the block's length and its height and depth."""
    p, pos, _ = scenes.double_dam_break(seed=0)
    assert pos.shape == (8_388_608, 3) and p.max_neighbors == 64
    assert tuple(p.boundary_max) == pytest.approx((25.6, 12.8, 12.8))
    left = pos[: pos.shape[0] // 2]
    right = pos[pos.shape[0] // 2:]
    assert left[:, 0].max() < 6.4 + 0.001 and right[:, 0].min() > 19.2 - 0.001
    assert pos[:, 1].max() < 12.8 and pos[:, 2].max() < 6.4 + 0.001
    del pos, left, right
    params, boxes, sc = scenes.batched_dam_breaks(num_scenes=256, seed=0)
    assert len(sc) == 256 and boxes.shape == (256, 3)
    counts = np.array([x.shape[0] for x, _ in sc])
    assert counts.min() >= 3000 and counts.max() <= 80000
    assert 5_000_000 < counts.sum() < 6_500_000 and params.max_particles_local == counts.sum()
    s = np.float32(scenes.SPACING)
    for (x, _), b in zip(sc, boxes):
        ext = x.max(0) - x.min(0) + s  # block extent: lattice span plus one spacing
        assert np.all(x >= 0) and np.all(x < b)
        assert b[0] == pytest.approx(2 * ext[0], rel=0.02)
        assert b[1] == pytest.approx(ext[1], rel=0.02) and b[2] == pytest.approx(ext[2], rel=0.02)
        assert 0.4 <= ext[0] / ext[1] <= 2.1 and 0.4 <= ext[2] / ext[1] <= 2.1


def _header_text():
    return re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)


def test_python_constants_match_the_header(pbf):
    """Every integer #define of include/pbf.h that the package mirrors has the
    header's value, and every option bit is mirrored."""
    defs = dict(re.findall(r"^#define (PBF_[A-Z0-9_]+) (\d+)u?\s*$", _header_text(), flags=re.M))
    assert any(k.startswith("PBF_OPT_") for k in defs)
    for name, value in defs.items():
        if name.startswith("PBF_OPT_"):
            assert hasattr(pbf, name), name
        if hasattr(pbf, name):
            assert getattr(pbf, name) == int(value), name
    bits = [int(v) for k, v in defs.items() if k.startswith("PBF_OPT_")]
    assert len(set(bits)) == len(bits) and all(b & (b - 1) == 0 for b in bits)  # distinct single bits


@pytest.mark.parametrize("c_name,py_name", [("pbf_stats", "Stats"), ("pbf_solve_info", "SolveInfo"),
                                            ("pbf_solve_record", "SolveRecord"), ("pbf_tuning", "Tuning"),
                                            ("pbf_layout", "Layout")])
def test_python_structs_match_the_header(pbf, c_name, py_name):
    """The ctypes mirrors list the C structs' fields in the same order."""
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (c_name, c_name), _header_text(), flags=re.S).group(1)
    c_fields = re.findall(r"\b([a-z_][a-z0-9_]*)(?:\[[^\]]*\])?\s*;", body)
    py_fields = [f[0] for f in getattr(pbf, py_name)._fields_]
    assert c_fields == py_fields, (c_fields, py_fields)
