// particles.hpp -- C++ host-side mirror of sim::Particles<float,3> over the C ABI.
//
// Simulation/Source/particles.h:41-504 declares the particle state and one
// method per PBF pass; Simulation/Source/neighbors.h:61-285 the neighbour
// grid.  This header keeps those method names and their IndexSpan arguments
// and forwards each call to include/pbf.h, so host code written against the
// reference (Simulation/Source/main.cpp) runs on the MI355X library by
// changing the type it constructs.  Failures throw std::runtime_error, as the
// reference's Array and ManagedAllocation do (array.h:55,129).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "pbf.h"

namespace sim {
namespace mi355x {

// Vec<size_t,2>{begin, end} (vec.h:706)
struct IndexSpan {
  std::size_t begin;
  std::size_t end;
};

inline void check(int rc) {
  if (rc != PBF_OK) throw std::runtime_error(pbf_last_error());
}

inline pbf_index_span c_span(IndexSpan s) { return pbf_index_span{s.begin, s.end}; }

class Particles {
 public:
  explicit Particles(const pbf_params& parameters) : params_(parameters) { check(pbf_create(&params_, &ctx_)); }
  // independent scenes in one context: boxes holds 3 extents per scene
  Particles(const pbf_params& parameters, const std::vector<float>& boxes) : params_(parameters) {
    if (boxes.size() % 3 != 0 || boxes.empty()) throw std::runtime_error("boxes: 3 extents per scene");
    check(pbf_create_batched(&params_, (uint32_t)(boxes.size() / 3), boxes.data(), &ctx_));
  }
  ~Particles() { pbf_destroy(ctx_); }
  Particles(const Particles&) = delete;
  Particles& operator=(const Particles&) = delete;

  // particles.h:76-86
  std::size_t local_count() const { return pbf_local_count(ctx_); }
  std::size_t available() const { return pbf_available(ctx_); }
  std::size_t max_local_count() const { return pbf_max_local_count(ctx_); }

  // particles.h:184-199 (3 floats per particle: Vec<float,3>)
  void add(const float* positions, const float* position_stars, const float* velocities, std::size_t count) {
    check(pbf_add(ctx_, positions, position_stars, velocities, count));
  }
  void remove(std::size_t count) { check(pbf_remove(ctx_, count)); }
  void add_to_scene(uint32_t scene, const float* positions, const float* position_stars, const float* velocities,
                    std::size_t count) {
    check(pbf_add_to_scene(ctx_, scene, positions, position_stars, velocities, count));
  }

  // renderer re-broadcast (Simulation/Source/main.cpp:62-63)
  void set_parameters(const pbf_params& parameters) {
    params_ = parameters;
    check(pbf_set_params(ctx_, &params_));
  }
  const pbf_params& parameters() const { return params_; }

  // particles.h:256-487
  void find_neighbors(IndexSpan to_bin_span, IndexSpan to_fill_span) {
    check(pbf_find_neighbors(ctx_, c_span(to_bin_span), c_span(to_fill_span)));
  }
  void apply_external_forces(IndexSpan s) { check(pbf_apply_external_forces(ctx_, c_span(s))); }
  void predict_positions(IndexSpan s) { check(pbf_predict_positions(ctx_, c_span(s))); }
  void compute_densities(IndexSpan s) { check(pbf_compute_densities(ctx_, c_span(s))); }
  void compute_pressure_lambdas(IndexSpan s) { check(pbf_compute_pressure_lambdas(ctx_, c_span(s))); }
  void compute_pressure_dps(IndexSpan s, int substep) { check(pbf_compute_pressure_dps(ctx_, c_span(s), substep)); }
  void update_position_stars(IndexSpan s) { check(pbf_update_position_stars(ctx_, c_span(s))); }
  void update_velocities(IndexSpan s) { check(pbf_update_velocities(ctx_, c_span(s))); }
  void update_positions(IndexSpan s) { check(pbf_update_positions(ctx_, c_span(s))); }
  void apply_surface_tension(IndexSpan color_field_span, IndexSpan surface_tension_span) {
    check(pbf_apply_surface_tension(ctx_, c_span(color_field_span), c_span(surface_tension_span)));
  }
  void apply_viscosity(IndexSpan s) { check(pbf_apply_viscosity(ctx_, c_span(s))); }
  void compute_vorticity(IndexSpan s) { check(pbf_compute_vorticity(ctx_, c_span(s))); }
  void apply_vorticity(IndexSpan s) { check(pbf_apply_vorticity(ctx_, c_span(s))); }

  // the whole main.cpp step body, fused; through a replayed HIP graph
  void step(unsigned nsteps) { check(pbf_step(ctx_, nsteps)); }
  void step_graph(unsigned nsteps) { check(pbf_step_graph(ctx_, nsteps)); }
  // one step iterating the pressure solve to a density-error tolerance
  pbf_solve_info solve(float tolerance, unsigned max_iterations, int warm_start = PBF_WARM_NONE) {
    check(pbf_solve(ctx_, tolerance, max_iterations, warm_start));
    pbf_solve_info info;
    check(pbf_get_solve_info(ctx_, &info));
    return info;
  }
  // nsteps such steps replayed from captured HIP graphs (no host round trip),
  // then their per-step records
  std::vector<pbf_solve_record> solve_graph(float tolerance, unsigned max_iterations, int warm_start,
                                            unsigned nsteps) {
    check(pbf_solve_graph(ctx_, tolerance, max_iterations, warm_start, nsteps));
    std::vector<pbf_solve_record> log(nsteps < PBF_SOLVE_LOG_SIZE ? nsteps : PBF_SOLVE_LOG_SIZE);
    uint32_t got = 0;
    check(pbf_get_solve_log(ctx_, log.data(), (uint32_t)log.size(), &got));
    log.resize(got);
    return log;
  }
  void set_warm_lambdas(const float* lambdas) { check(pbf_set_warm_lambdas(ctx_, lambdas)); }
  void set_warm_model(pbf_warm_model* model) { check(pbf_set_warm_model(ctx_, model)); }
  pbf_tuning autotune(unsigned steps) {
    pbf_tuning t;
    double ms = 0.0;
    check(pbf_autotune(ctx_, steps, &t, &ms));
    return t;
  }
  // bin spacing and list cap for speed (changes the list layout against the reference's)
  pbf_layout autotune_layout(unsigned steps) {
    pbf_layout l;
    check(pbf_autotune_layout(ctx_, steps, &l));
    check(pbf_get_params(ctx_, &params_));
    return l;
  }
  void sync() { check(pbf_sync(ctx_)); }

  // array getters (particles.h:91-141), original order
  std::vector<float> positions() { return get3(pbf_get_positions); }
  std::vector<float> position_stars() { return get3(pbf_get_position_stars); }
  std::vector<float> velocities() { return get3(pbf_get_velocities); }
  std::vector<float> densities() {
    std::vector<float> out(local_count());
    check(pbf_get_densities(ctx_, out.data()));
    return out;
  }

  pbf_stats stats() {
    pbf_stats s;
    check(pbf_get_stats(ctx_, &s));
    return s;
  }

  // AdiosWriter::write_particles record (adios_writer.h:78-114) as raw bytes
  void write_positions(const std::string& path) { check(pbf_append_positions_record(ctx_, path.c_str())); }

  pbf_context* handle() { return ctx_; }

 private:
  std::vector<float> get3(int (*getter)(pbf_context*, float*)) {
    std::vector<float> out(local_count() * 3);
    check(getter(ctx_, out.data()));
    return out;
  }

  pbf_params params_;
  pbf_context* ctx_ = nullptr;
};

}  // namespace mi355x
}  // namespace sim
