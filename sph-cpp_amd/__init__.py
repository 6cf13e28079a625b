"""Python mirror of the PBF library's C ABI (include/pbf.h).

The class :class:`Particles` keeps the method names and argument meaning of
``sim::Particles<float,3>`` (Simulation/Source/particles.h:41-504) and the
neighbour introspection of ``sim::Neighbors`` (Simulation/Source/neighbors.h),
so parity tests read like the reference's own tests
(Simulation/Tests/Serial/particle_test.cpp, neighbor_test.cpp).  Every call
goes through ``libpbf.so``; there is no other code path.  If the library is
missing, :func:`lib` raises.

Import with ``importlib.import_module("sph-cpp_amd")``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbf.so")

PBF_OK = 0
PBF_ERR_ARG = 1
PBF_ERR_CAPACITY = 2
PBF_ERR_HIP = 3
PBF_ERR_STATE = 4
PBF_ERR_UNSUPPORTED = 5

PBF_OPT_COUNT_OVERFLOW = 1


class PBFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pbf error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """pbf_params: Parameters<float,3> fields used by the step (parameters.h:550-573)."""

    _fields_ = [
        ("max_particles_local", C.c_uint64),
        ("solve_step_count", C.c_uint32),
        ("max_neighbors", C.c_uint32),
        ("particle_rest_spacing", C.c_float),
        ("particle_radius", C.c_float),
        ("smoothing_radius", C.c_float),
        ("neighbor_bin_spacing", C.c_float),
        ("rest_density", C.c_float),
        ("rest_mass", C.c_float),
        ("gravity", C.c_float),
        ("gamma", C.c_float),
        ("lambda_epsilon", C.c_float),
        ("k_stiff", C.c_float),
        ("visc_c", C.c_float),
        ("time_step", C.c_float),
        ("max_speed", C.c_float),
        ("vorticity_coef", C.c_float),
        ("boundary_min", C.c_float * 3),
        ("boundary_max", C.c_float * 3),
        ("mover_center", C.c_float * 3),
        ("options", C.c_uint32),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class IndexSpan(C.Structure):
    """IndexSpan = Vec<size_t,2>{begin, end} (vec.h:706)."""

    _fields_ = [("begin", C.c_uint64), ("end", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [
        ("neighbor_overflow", C.c_uint64),
        ("outside_grid", C.c_uint64),
        ("density_hacks", C.c_uint64),
        ("iterations", C.c_uint64),
        ("max_density_error", C.c_double),
        ("tile_listed", C.c_uint64),
        ("persistent_grid", C.c_uint64),
        ("persistent_capacity", C.c_uint64),
    ]


PBF_MAX_SOLVE_ITERATIONS = 256
PBF_WARM_NONE, PBF_WARM_GIVEN, PBF_WARM_PREVIOUS, PBF_WARM_LEARNED = 0, 1, 2, 3
PBF_OPT_SOLVE_MEAN_ERROR = 2
PBF_OPT_KEEP_WARM_FEATURES = 4
PBF_OPT_FAST_MATH = 8
PBF_OPT_WARM_STRICT = 16
PBF_OPT_FULL_LISTS = 32
PBF_OPT_LANE_LISTS = 64
PBF_OPT_PERSISTENT = 128
PBF_OPT_WARM_RETRY = 256
PBF_WARM_FEATURES = 8
PBF_WARM_HIDDEN = 64
PBF_WARM_PARAMS = 4801
PBF_WARM_LAMBDA_SCALE = 1e-4


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("heavy_budget", C.c_uint32), ("heavy_blocks", C.c_uint32)]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("warm_start_kept", C.c_int32),
        ("max_error_before_warm_start", C.c_float),
        ("mean_error_before_warm_start", C.c_float),
        ("max_error_history", C.c_float * PBF_MAX_SOLVE_ITERATIONS),
        ("mean_error_history", C.c_float * PBF_MAX_SOLVE_ITERATIONS),
    ]


class Layout(C.Structure):
    _fields_ = [("neighbor_bin_spacing", C.c_float), ("max_neighbors", C.c_uint32), ("ms_per_step", C.c_double),
                ("longest_list", C.c_uint32)]


class SolveRecord(C.Structure):
    _fields_ = [
        ("index", C.c_uint64),
        ("iterations", C.c_uint32),
        ("warm_start_kept", C.c_int32),
        ("max_error", C.c_float),
        ("mean_error", C.c_float),
        ("max_error_before_warm_start", C.c_float),
        ("mean_error_before_warm_start", C.c_float),
    ]


PBF_SOLVE_LOG_SIZE = 4096

_lib = None

_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U64 = C.POINTER(C.c_uint64)
_I32 = C.POINTER(C.c_int32)


def _sig(lib, name, args, res=C.c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res


def lib():
    """Load libpbf.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python sph-cpp_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    _sig(L, "pbf_last_error", [], C.c_char_p)
    _sig(L, "pbf_params_derive", [C.POINTER(Params)])
    _sig(L, "pbf_create", [C.POINTER(Params), C.POINTER(_P)])
    _sig(L, "pbf_destroy", [_P])
    _sig(L, "pbf_set_params", [_P, C.POINTER(Params)])
    _sig(L, "pbf_get_params", [_P, C.POINTER(Params)])
    _sig(L, "pbf_add", [_P, _F, _F, _F, C.c_uint64])
    _sig(L, "pbf_remove", [_P, C.c_uint64])
    for nm in ("pbf_local_count", "pbf_available", "pbf_max_local_count", "pbf_num_cells"):
        _sig(L, nm, [_P], C.c_uint64)
    for nm in ("positions", "position_stars", "velocities", "densities", "lambdas", "scratch"):
        _sig(L, "pbf_get_" + nm, [_P, _F])
        _sig(L, "pbf_set_" + nm, [_P, _F])
    for nm in ("apply_external_forces", "predict_positions", "compute_densities", "compute_pressure_lambdas",
               "update_position_stars", "update_velocities", "apply_viscosity", "compute_vorticity",
               "apply_vorticity", "update_positions"):
        _sig(L, "pbf_" + nm, [_P, IndexSpan])
    _sig(L, "pbf_compute_pressure_dps", [_P, IndexSpan, C.c_int])
    _sig(L, "pbf_find_neighbors", [_P, IndexSpan, IndexSpan])
    _sig(L, "pbf_apply_surface_tension", [_P, IndexSpan, IndexSpan])
    _sig(L, "pbf_step", [_P, C.c_uint32])
    _sig(L, "pbf_solve", [_P, C.c_float, C.c_uint32, C.c_int])
    _sig(L, "pbf_solve_graph", [_P, C.c_float, C.c_uint32, C.c_int, C.c_uint32])
    _sig(L, "pbf_get_solve_log", [_P, C.POINTER(SolveRecord), C.c_uint32, C.POINTER(C.c_uint32)])
    _sig(L, "pbf_step_graph", [_P, C.c_uint32])
    _sig(L, "pbf_set_tuning", [_P, C.POINTER(Tuning)])
    _sig(L, "pbf_get_tuning", [_P, C.POINTER(Tuning)])
    _sig(L, "pbf_autotune", [_P, C.c_uint32, C.POINTER(Tuning), C.POINTER(C.c_double)])
    _sig(L, "pbf_autotune_layout", [_P, C.c_uint32, C.POINTER(Layout)])
    _sig(L, "pbf_create_batched", [C.POINTER(Params), C.c_uint32, _F, C.POINTER(_P)])
    _sig(L, "pbf_add_to_scene", [_P, C.c_uint32, _F, _F, _F, C.c_uint64])
    L.pbf_num_scenes.argtypes = [_P]
    L.pbf_num_scenes.restype = C.c_uint32
    _sig(L, "pbf_set_warm_lambdas", [_P, _F])
    _sig(L, "pbf_get_lambda_sums", [_P, _F])
    _sig(L, "pbf_get_solve_info", [_P, C.POINTER(SolveInfo)])
    _sig(L, "pbf_get_warm_features", [_P, _F])
    _sig(L, "pbf_set_lambda_sums", [_P, _F])
    _sig(L, "pbf_set_warm_model", [_P, _P])
    _sig(L, "pbf_set_warm_gain", [_P, C.c_float])
    _sig(L, "pbf_warm_model_create", [C.c_uint64, C.POINTER(_P)])
    _sig(L, "pbf_warm_model_destroy", [_P])
    _sig(L, "pbf_warm_model_init", [_P, C.c_uint64])
    _sig(L, "pbf_warm_model_set_params", [_P, _F])
    _sig(L, "pbf_warm_model_get_params", [_P, _F])
    _sig(L, "pbf_warm_model_add_samples", [_P, _F, _F, C.c_uint64])
    _sig(L, "pbf_warm_model_add_from_context", [_P, _P, C.c_uint32])
    _sig(L, "pbf_warm_model_clear_samples", [_P])
    L.pbf_warm_model_sample_count.argtypes = [_P]
    L.pbf_warm_model_sample_count.restype = C.c_uint64
    _sig(L, "pbf_warm_model_gradient", [_P, C.c_uint64, C.c_uint32, _F, _F])
    _sig(L, "pbf_warm_model_adam_step", [_P, _F, C.c_float])
    _sig(L, "pbf_warm_model_train", [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _F])
    _sig(L, "pbf_warm_model_predict", [_P, _F, C.c_uint64, _F])
    _sig(L, "pbf_check_arithmetic", [C.c_uint64, C.c_uint64, _U64])
    _sig(L, "pbf_sync", [_P])
    _sig(L, "pbf_get_stats", [_P, C.POINTER(Stats)])
    _sig(L, "pbf_grid_dims", [_P, _U64])
    _sig(L, "pbf_get_bin_ids", [_P, _U64])
    _sig(L, "pbf_get_particle_ids", [_P, _U64])
    _sig(L, "pbf_get_bin_bounds", [_P, _U64, _U64])
    _sig(L, "pbf_get_neighbor_lists", [_P, _U64, _I32])
    _sig(L, "pbf_sort_by_key_u32", [_P, _P, C.c_uint64, C.c_uint32])
    _sig(L, "pbf_sort_by_key_u32_noalloc", [_P, _P, C.c_uint64, C.c_uint32])
    _sig(L, "pbf_lower_bound_counting", [_P, C.c_uint64, IndexSpan, _P])
    _sig(L, "pbf_upper_bound_counting", [_P, C.c_uint64, IndexSpan, _P])
    _sig(L, "pbf_device_alloc", [C.c_uint64, C.POINTER(_P)])
    _sig(L, "pbf_device_free", [_P])
    _sig(L, "pbf_memcpy_to_device", [_P, _P, C.c_uint64])
    _sig(L, "pbf_memcpy_to_host", [_P, _P, C.c_uint64])
    _sig(L, "pbf_memcpy_device", [_P, _P, C.c_uint64])
    _sig(L, "pbf_append_positions_record", [_P, C.c_char_p])
    _sig(L, "pbf_profile_kernels", [_P, C.c_char_p])
    _sig(L, "pbf_kernel_time", [_P, C.c_char_p, _U64, C.POINTER(C.c_double)])
    _sig(L, "pbf_region_begin", [_P])
    _sig(L, "pbf_region_end", [_P, C.POINTER(C.c_double)])
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != PBF_OK:
        raise PBFError(rc, lib().pbf_last_error().decode())


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_F)


def derive(p: Params) -> Params:
    """Parameters::derive_from_input (parameters.h:123-148) through the library."""
    _check(lib().pbf_params_derive(C.byref(p)))
    return p


def span(begin: int, end: int) -> IndexSpan:
    return IndexSpan(int(begin), int(end))


class Particles:
    """sim::Particles<float,3> on the device (particles.h:41-504)."""

    def __init__(self, params: Params, boxes=None):
        """One domain; with ``boxes`` (num_scenes x 3 box extents) a batched
        context of independent scenes (pbf_create_batched)."""
        L = lib()
        self._ctx = _P()
        self.params = params.copy()
        if boxes is None:
            _check(L.pbf_create(C.byref(self.params), C.byref(self._ctx)))
        else:
            b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 3)
            _check(L.pbf_create_batched(C.byref(self.params), b.shape[0], _fptr(b), C.byref(self._ctx)))

    def close(self) -> None:
        if self._ctx:
            lib().pbf_destroy(self._ctx)
            self._ctx = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- storage (particles.h:76-199) ----
    def local_count(self) -> int:
        return int(lib().pbf_local_count(self._ctx))

    def available(self) -> int:
        return int(lib().pbf_available(self._ctx))

    def max_local_count(self) -> int:
        return int(lib().pbf_max_local_count(self._ctx))

    def add(self, positions, position_stars=None, velocities=None) -> None:
        x = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        xs = x if position_stars is None else np.ascontiguousarray(position_stars, dtype=np.float32).reshape(-1, 3)
        v = np.zeros_like(x) if velocities is None else np.ascontiguousarray(velocities, np.float32).reshape(-1, 3)
        _check(lib().pbf_add(self._ctx, _fptr(x), _fptr(xs), _fptr(v), x.shape[0]))

    def add_to_scene(self, scene: int, positions, position_stars=None, velocities=None) -> None:
        x = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        xs = x if position_stars is None else np.ascontiguousarray(position_stars, dtype=np.float32).reshape(-1, 3)
        v = np.zeros_like(x) if velocities is None else np.ascontiguousarray(velocities, np.float32).reshape(-1, 3)
        _check(lib().pbf_add_to_scene(self._ctx, int(scene), _fptr(x), _fptr(xs), _fptr(v), x.shape[0]))

    def num_scenes(self) -> int:
        return int(lib().pbf_num_scenes(self._ctx))

    def remove(self, count: int) -> None:
        _check(lib().pbf_remove(self._ctx, int(count)))

    def set_params(self, params: Params) -> None:
        p = params.copy()
        _check(lib().pbf_set_params(self._ctx, C.byref(p)))
        self.params = p

    def _get(self, name: str, comps: int) -> np.ndarray:
        n = self.local_count()
        out = np.empty((n, comps) if comps > 1 else (n,), dtype=np.float32)
        _check(getattr(lib(), "pbf_get_" + name)(self._ctx, _fptr(out)))
        return out

    def _set(self, name: str, a, comps: int) -> None:
        n = self.local_count()
        arr = np.ascontiguousarray(a, dtype=np.float32).reshape((n, comps) if comps > 1 else (n,))
        _check(getattr(lib(), "pbf_set_" + name)(self._ctx, _fptr(arr)))

    def positions(self):
        return self._get("positions", 3)

    def position_stars(self):
        return self._get("position_stars", 3)

    def velocities(self):
        return self._get("velocities", 3)

    def densities(self):
        return self._get("densities", 1)

    def lambdas(self):
        return self._get("lambdas", 1)

    def scratch(self):
        return self._get("scratch", 3)

    def set_positions(self, a):
        self._set("positions", a, 3)

    def set_position_stars(self, a):
        self._set("position_stars", a, 3)

    def set_velocities(self, a):
        self._set("velocities", a, 3)

    def set_densities(self, a):
        self._set("densities", a, 1)

    def set_lambdas(self, a):
        self._set("lambdas", a, 1)

    def set_scratch(self, a):
        self._set("scratch", a, 3)

    # ---- passes (particles.h:256-487) ----
    def _all(self, s):
        return span(0, self.local_count()) if s is None else s

    def find_neighbors(self, to_bin_span=None, to_fill_span=None):
        _check(lib().pbf_find_neighbors(self._ctx, self._all(to_bin_span), self._all(to_fill_span)))

    def apply_external_forces(self, s=None):
        _check(lib().pbf_apply_external_forces(self._ctx, self._all(s)))

    def predict_positions(self, s=None):
        _check(lib().pbf_predict_positions(self._ctx, self._all(s)))

    def compute_densities(self, s=None):
        _check(lib().pbf_compute_densities(self._ctx, self._all(s)))

    def compute_pressure_lambdas(self, s=None):
        _check(lib().pbf_compute_pressure_lambdas(self._ctx, self._all(s)))

    def compute_pressure_dps(self, s=None, substep: int = 0):
        _check(lib().pbf_compute_pressure_dps(self._ctx, self._all(s), int(substep)))

    def update_position_stars(self, s=None):
        _check(lib().pbf_update_position_stars(self._ctx, self._all(s)))

    def update_velocities(self, s=None):
        _check(lib().pbf_update_velocities(self._ctx, self._all(s)))

    def update_positions(self, s=None):
        _check(lib().pbf_update_positions(self._ctx, self._all(s)))

    def apply_surface_tension(self, color_field_span=None, surface_tension_span=None):
        _check(lib().pbf_apply_surface_tension(self._ctx, self._all(color_field_span),
                                               self._all(surface_tension_span)))

    def apply_viscosity(self, s=None):
        _check(lib().pbf_apply_viscosity(self._ctx, self._all(s)))

    def compute_vorticity(self, s=None):
        _check(lib().pbf_compute_vorticity(self._ctx, self._all(s)))

    def apply_vorticity(self, s=None):
        _check(lib().pbf_apply_vorticity(self._ctx, self._all(s)))

    # ---- whole step and sync ----
    def step(self, nsteps: int = 1) -> None:
        """Simulation/Source/main.cpp:71-124 on one domain, nsteps times (fused kernels)."""
        _check(lib().pbf_step(self._ctx, int(nsteps)))

    def solve(self, tolerance: float, max_iterations: int, warm_start: int = PBF_WARM_NONE) -> None:
        """One step with an iterate-to-tolerance pressure solve (pbf_solve)."""
        _check(lib().pbf_solve(self._ctx, float(tolerance), int(max_iterations), int(warm_start)))

    def solve_graph(self, tolerance: float, max_iterations: int, warm_start: int = PBF_WARM_NONE,
                    nsteps: int = 1) -> None:
        """nsteps iterate-to-tolerance steps replayed from a HIP graph (pbf_solve_graph)."""
        _check(lib().pbf_solve_graph(self._ctx, float(tolerance), int(max_iterations), int(warm_start),
                                     int(nsteps)))

    def solve_log(self, count: int) -> list:
        """The last `count` per-solve records (pbf_get_solve_log), oldest first."""
        count = min(int(count), PBF_SOLVE_LOG_SIZE)
        recs = (SolveRecord * max(1, count))()
        got = C.c_uint32(0)
        _check(lib().pbf_get_solve_log(self._ctx, recs, count, C.byref(got)))
        return [{"index": r.index, "iterations": r.iterations, "warm_start_kept": bool(r.warm_start_kept),
                 "warm_start_damped": r.warm_start_kept == 2,
                 "max_error": r.max_error, "mean_error": r.mean_error,
                 "max_error_before_warm_start": r.max_error_before_warm_start,
                 "mean_error_before_warm_start": r.mean_error_before_warm_start} for r in recs[:got.value]]

    def set_warm_lambdas(self, lambdas) -> None:
        a = np.ascontiguousarray(lambdas, dtype=np.float32).reshape(-1)
        if a.shape[0] != self.local_count():
            raise ValueError("one multiplier per local particle")
        _check(lib().pbf_set_warm_lambdas(self._ctx, a.ctypes.data_as(_F)))

    def lambda_sums(self) -> np.ndarray:
        out = np.empty(self.local_count(), dtype=np.float32)
        _check(lib().pbf_get_lambda_sums(self._ctx, out.ctypes.data_as(_F)))
        return out

    def set_lambda_sums(self, sums) -> None:
        a = np.ascontiguousarray(sums, dtype=np.float32).reshape(-1)
        if a.shape[0] != self.local_count():
            raise ValueError("one sum per local particle")
        _check(lib().pbf_set_lambda_sums(self._ctx, a.ctypes.data_as(_F)))

    def solve_info(self) -> dict:
        si = SolveInfo()
        _check(lib().pbf_get_solve_info(self._ctx, C.byref(si)))
        k = si.iterations
        return {"iterations": k, "warm_start_kept": bool(si.warm_start_kept),
                "warm_start_damped": si.warm_start_kept == 2,
                "max_error_before_warm_start": si.max_error_before_warm_start,
                "mean_error_before_warm_start": si.mean_error_before_warm_start,
                "max_error_history": np.array(si.max_error_history[:k], dtype=np.float32),
                "mean_error_history": np.array(si.mean_error_history[:k], dtype=np.float32)}

    def warm_features(self) -> np.ndarray:
        out = np.empty((self.local_count(), PBF_WARM_FEATURES), dtype=np.float32)
        _check(lib().pbf_get_warm_features(self._ctx, out.ctypes.data_as(_F)))
        return out

    def set_warm_model(self, model: "WarmModel", gain: float = 1.0) -> None:
        _check(lib().pbf_set_warm_model(self._ctx, model._h))
        _check(lib().pbf_set_warm_gain(self._ctx, float(gain)))

    def set_tuning(self, block_size: int = 256, heavy_budget: int = 768, heavy_blocks: int = 8192) -> None:
        t = Tuning(block_size, heavy_budget, heavy_blocks)
        _check(lib().pbf_set_tuning(self._ctx, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().pbf_get_tuning(self._ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def autotune(self, steps: int = 3) -> dict:
        """pbf_autotune: time the launch settings from the current state, keep the fastest."""
        t = Tuning()
        ms = C.c_double(0.0)
        _check(lib().pbf_autotune(self._ctx, int(steps), C.byref(t), C.byref(ms)))
        return {**{k: getattr(t, k) for k, _ in Tuning._fields_}, "ms_per_step": ms.value}

    def autotune_layout(self, steps: int = 3) -> dict:
        """pbf_autotune_layout: pick the bin spacing and list cap (changes the
        list layout against the reference's; see include/pbf.h)."""
        lay = Layout()
        _check(lib().pbf_autotune_layout(self._ctx, int(steps), C.byref(lay)))
        p = Params()
        _check(lib().pbf_get_params(self._ctx, C.byref(p)))
        self.params = p
        return {k: getattr(lay, k) for k, _ in Layout._fields_}

    def step_graph(self, nsteps: int = 1) -> None:
        """pbf_step through a captured HIP graph of two steps."""
        _check(lib().pbf_step_graph(self._ctx, int(nsteps)))

    def sync(self) -> None:
        _check(lib().pbf_sync(self._ctx))

    def stats(self) -> dict:
        s = Stats()
        _check(lib().pbf_get_stats(self._ctx, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    # ---- Neighbors introspection (neighbors.h:250-284) ----
    def bin_dimensions(self):
        d = (C.c_uint64 * 3)()
        _check(lib().pbf_grid_dims(self._ctx, d))
        return tuple(int(x) for x in d)

    def num_cells(self) -> int:
        return int(lib().pbf_num_cells(self._ctx))

    def bin_ids(self) -> np.ndarray:
        out = np.empty(self.local_count(), dtype=np.uint64)
        _check(lib().pbf_get_bin_ids(self._ctx, out.ctypes.data_as(_U64)))
        return out

    def particle_ids(self) -> np.ndarray:
        out = np.empty(self.local_count(), dtype=np.uint64)
        _check(lib().pbf_get_particle_ids(self._ctx, out.ctypes.data_as(_U64)))
        return out

    def bin_bounds(self):
        nc = self.num_cells()
        b = np.empty(nc, dtype=np.uint64)
        e = np.empty(nc, dtype=np.uint64)
        _check(lib().pbf_get_bin_bounds(self._ctx, b.ctypes.data_as(_U64), e.ctypes.data_as(_U64)))
        return b, e

    def neighbor_lists(self):
        n = self.local_count()
        cap = int(self.params.max_neighbors)
        lists = np.zeros((n, cap), dtype=np.uint64)
        counts = np.zeros(n, dtype=np.int32)
        _check(lib().pbf_get_neighbor_lists(self._ctx, lists.ctypes.data_as(_U64), counts.ctypes.data_as(_I32)))
        return lists, counts

    def append_positions_record(self, path: str) -> None:
        """AdiosWriter::write_particles record layout (adios_writer.h:78-114), raw."""
        _check(lib().pbf_append_positions_record(self._ctx, path.encode()))

    # ---- measurement ----
    def profile_kernels(self, names) -> None:
        if not isinstance(names, str):
            names = ",".join(names)
        _check(lib().pbf_profile_kernels(self._ctx, names.encode()))

    def kernel_time(self, name: str):
        n = C.c_uint64()
        ms = C.c_double()
        _check(lib().pbf_kernel_time(self._ctx, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def region_begin(self) -> None:
        _check(lib().pbf_region_begin(self._ctx))

    def region_end(self) -> float:
        ms = C.c_double()
        _check(lib().pbf_region_end(self._ctx, C.byref(ms)))
        return float(ms.value)


# ---- sim::algorithms primitives on device memory ----
class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.ptr = _P()
        self.nbytes = int(nbytes)
        _check(lib().pbf_device_alloc(self.nbytes, C.byref(self.ptr)))

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _check(lib().pbf_memcpy_to_device(b.ptr, a.ctypes.data_as(_P), a.nbytes))
        return b

    def to_array(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(lib().pbf_memcpy_to_host(out.ctypes.data_as(_P), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().pbf_device_free(self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def check_arithmetic(samples: int, seed: int = 0) -> dict:
    """pbf_check_arithmetic: the library's divide and square root against the
    plain operations on the device."""
    out = np.zeros(5, dtype=np.uint64)
    _check(lib().pbf_check_arithmetic(int(samples), int(seed), out.ctypes.data_as(_U64)))
    return {"divide_mismatches": int(out[0]), "sqrt_mismatches": int(out[1]),
            "divide_fast": int(out[2]), "sqrt_fast": int(out[3]), "divide_mismatches_shaped": int(out[4])}


NOALLOC_MAX_PAIRS = 1 << 20  # PBF_NOALLOC_MAX_PAIRS


def sort_by_key_u32(keys: np.ndarray, values: np.ndarray, key_bits: int = 32, no_alloc: bool = False):
    """sim::algorithms::sort_by_key (sim_algorithms.h:57-61): stable, on the device.
    no_alloc=True calls pbf_sort_by_key_u32_noalloc (fixed work space, at most
    NOALLOC_MAX_PAIRS pairs)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    values = np.ascontiguousarray(values, dtype=np.uint32)
    kb, vb = DeviceBuffer.from_array(keys), DeviceBuffer.from_array(values)
    f = lib().pbf_sort_by_key_u32_noalloc if no_alloc else lib().pbf_sort_by_key_u32
    _check(f(kb.ptr, vb.ptr, keys.size, int(key_bits)))
    return kb.to_array(np.uint32, keys.size), vb.to_array(np.uint32, keys.size)


def bound_counting(sorted_keys: np.ndarray, begin: int, end: int, upper: bool) -> np.ndarray:
    """sim::algorithms::lower_bound / upper_bound over a counting range (sim_algorithms.h:69-99)."""
    sk = np.ascontiguousarray(sorted_keys, dtype=np.uint32)
    kb = DeviceBuffer.from_array(sk)
    ob = DeviceBuffer(max(1, (end - begin) * 4))
    f = lib().pbf_upper_bound_counting if upper else lib().pbf_lower_bound_counting
    _check(f(kb.ptr, sk.size, span(begin, end), ob.ptr))
    return ob.to_array(np.uint32, end - begin)


class WarmModel:
    """Learned warm start: the 8-64-64-1 network of pbf_warm_model_* (training
    and inference on the GPU).  Targets and outputs are Lambda / PBF_WARM_LAMBDA_SCALE."""

    def __init__(self, max_samples: int):
        h = C.c_void_p()
        _check(lib().pbf_warm_model_create(int(max_samples), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            lib().pbf_warm_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, seed: int = 0) -> None:
        _check(lib().pbf_warm_model_init(self._h, int(seed)))

    def set_params(self, params) -> None:
        a = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        if a.shape[0] != PBF_WARM_PARAMS:
            raise ValueError(f"{PBF_WARM_PARAMS} parameters expected")
        _check(lib().pbf_warm_model_set_params(self._h, a.ctypes.data_as(_F)))

    def params(self) -> np.ndarray:
        out = np.empty(PBF_WARM_PARAMS, dtype=np.float32)
        _check(lib().pbf_warm_model_get_params(self._h, out.ctypes.data_as(_F)))
        return out

    def add_samples(self, features, targets) -> None:
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PBF_WARM_FEATURES)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        if f.shape[0] != t.shape[0]:
            raise ValueError("one target per feature row")
        _check(lib().pbf_warm_model_add_samples(self._h, f.ctypes.data_as(_F), t.ctypes.data_as(_F), t.shape[0]))

    def add_from(self, particles: "Particles", stride: int = 1) -> None:
        _check(lib().pbf_warm_model_add_from_context(self._h, particles._ctx, int(stride)))

    def clear_samples(self) -> None:
        _check(lib().pbf_warm_model_clear_samples(self._h))

    def sample_count(self) -> int:
        return int(lib().pbf_warm_model_sample_count(self._h))

    def gradient(self, first: int, count: int):
        g = np.empty(PBF_WARM_PARAMS, dtype=np.float32)
        loss = C.c_float(0.0)
        _check(lib().pbf_warm_model_gradient(self._h, int(first), int(count), g.ctypes.data_as(_F),
                                             C.byref(loss)))
        return g, float(loss.value)

    def adam_step(self, grad, learning_rate: float) -> None:
        g = np.ascontiguousarray(grad, dtype=np.float32)
        _check(lib().pbf_warm_model_adam_step(self._h, g.ctypes.data_as(_F), float(learning_rate)))

    def train(self, epochs: int, batch: int, learning_rate: float, seed: int = 0) -> float:
        loss = C.c_float(0.0)
        _check(lib().pbf_warm_model_train(self._h, int(epochs), int(batch), float(learning_rate), int(seed),
                                          C.byref(loss)))
        return float(loss.value)

    def predict(self, features) -> np.ndarray:
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PBF_WARM_FEATURES)
        out = np.empty(f.shape[0], dtype=np.float32)
        _check(lib().pbf_warm_model_predict(self._h, f.ctypes.data_as(_F), f.shape[0], out.ctypes.data_as(_F)))
        return out
