// sph_main.cpp -- single-domain step driver over the MI355X PBF library.
//
// Restates the compute loop of Simulation/Source/main.cpp:61-133 with the
// distributor collapsed to one domain (resident = local = all particles) and
// the renderer detached: the same passes in the same order, each a call
// through sim::mi355x::Particles (particles.hpp -> include/pbf.h).  With
// --fused the body is one pbf_step call per step instead.
//
// The initial fluid follows Particles::construct_fluid (particles.h:231-250):
// a lattice of points at cell centres, x fastest, plus a seeded uniform
// jitter.  Positions are appended every --every steps in the record layout
// of AdiosWriter (adios_writer.h:78-114).
//
//   sph_run [--lattice NXxNYxNZ] [--box X Y Z] [--steps N] [--iters I] [--cap K]
//           [--seed S] [--fused] [--graph] [--tolerance T] [--write FILE] [--every K]
// --graph replays the fused step through a HIP graph; --tolerance T runs
// each step as pbf_solve to a max density error of T (at most 64 iterations,
// warm-started from the previous step's multiplier sums).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "particles.hpp"

using sim::mi355x::IndexSpan;
using sim::mi355x::Particles;

namespace {

struct Options {
  int nx = 32, ny = 32, nz = 32;
  float box[3] = {3.2f, 1.6f, 1.6f};
  int steps = 10;
  int iters = 4;
  int cap = 60;
  unsigned seed = 0;
  bool fused = false;
  bool graph = false;
  float tolerance = -1.0f;
  std::string write;
  int every = 1;
};

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
      return argv[++i];
    };
    if (a == "--lattice") {
      if (std::sscanf(next(), "%dx%dx%d", &o.nx, &o.ny, &o.nz) != 3) throw std::runtime_error("bad --lattice");
    } else if (a == "--box") {
      for (int k = 0; k < 3; ++k) o.box[k] = std::strtof(next(), nullptr);
    } else if (a == "--steps") {
      o.steps = std::atoi(next());
    } else if (a == "--iters") {
      o.iters = std::atoi(next());
    } else if (a == "--cap") {
      o.cap = std::atoi(next());
    } else if (a == "--seed") {
      o.seed = (unsigned)std::strtoul(next(), nullptr, 10);
    } else if (a == "--fused") {
      o.fused = true;
    } else if (a == "--graph") {
      o.graph = true;
    } else if (a == "--tolerance") {
      o.tolerance = std::strtof(next(), nullptr);
    } else if (a == "--write") {
      o.write = next();
    } else if (a == "--every") {
      o.every = std::atoi(next());
    } else {
      throw std::runtime_error("unknown option " + a);
    }
  }
  return o;
}

// Parameters: derive_from_input (parameters.h:123-148), then the dam-break
// constants (rest density 1000, mass = rho0 * spacing^3).
pbf_params make_params(const Options& o, std::size_t n) {
  pbf_params p;
  std::memset(&p, 0, sizeof(p));
  p.max_particles_local = n;
  p.solve_step_count = (uint32_t)o.iters;
  p.max_neighbors = (uint32_t)o.cap;
  p.particle_rest_spacing = 0.05f;
  p.smoothing_radius = 0.1f;
  p.time_step = 1.0f / 120.0f;
  p.gravity = -9.8f;
  p.gamma = 0.0f;
  p.lambda_epsilon = 600.0f;
  p.k_stiff = 1.0f;
  p.visc_c = 0.01f;
  p.vorticity_coef = 0.0005f;
  for (int k = 0; k < 3; ++k) p.boundary_max[k] = o.box[k];
  p.mover_center[0] = -1.5f;
  p.mover_center[1] = -2.0f;
  p.mover_center[2] = -1.5f;
  sim::mi355x::check(pbf_params_derive(&p));
  p.rest_density = 1000.0f;
  p.rest_mass = 1000.0f * (0.05f * 0.05f * 0.05f);
  return p;
}

std::vector<float> make_fluid(const Options& o) {
  const float s = 0.05f;
  std::mt19937_64 rng(o.seed);
  std::uniform_real_distribution<float> jitter(-0.001f, 0.001f);
  std::vector<float> pos;
  pos.reserve((std::size_t)o.nx * o.ny * o.nz * 3);
  for (int z = 0; z < o.nz; ++z)
    for (int y = 0; y < o.ny; ++y)
      for (int x = 0; x < o.nx; ++x) {
        const int idx[3] = {x, y, z};
        for (int k = 0; k < 3; ++k) {
          float c = (float)idx[k] * s + 0.0f;  // int_coord * spacing + aabb.min
          c += s / 2.0f;
          pos.push_back(c);
        }
      }
  for (float& c : pos) c += jitter(rng);
  return pos;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Options o = parse(argc, argv);
    const std::vector<float> pos = make_fluid(o);
    const std::size_t n = pos.size() / 3;
    const std::vector<float> vel(pos.size(), 0.0f);
    Particles particles(make_params(o, n));
    particles.add(pos.data(), pos.data(), vel.data(), n);
    const IndexSpan all{0, n};  // resident_span() == local_span() on one domain
    const unsigned iters = particles.parameters().solve_step_count;

    const auto t0 = std::chrono::steady_clock::now();
    for (int frame = 0; frame < o.steps; ++frame) {
      if (o.tolerance >= 0.0f) {
        particles.solve(o.tolerance, 64, frame ? PBF_WARM_PREVIOUS : PBF_WARM_NONE);
      } else if (o.graph) {
        // the graph holds two steps: advance on even frames, covering the odd ones
        if (frame % 2 == 0) particles.step_graph(frame + 1 < o.steps ? 2 : 1);
      } else if (o.fused) {
        particles.step(1);
      } else {
        // Simulation/Source/main.cpp:71-124
        particles.apply_external_forces(all);
        particles.predict_positions(all);
        particles.find_neighbors(all, all);
        for (unsigned sub = 0; sub < iters; sub++) {
          particles.compute_densities(all);
          particles.compute_pressure_lambdas(all);
          particles.compute_pressure_dps(all, (int)sub);
          particles.update_position_stars(all);
        }
        particles.update_velocities(all);
        if (particles.parameters().gamma != 0.0f) particles.apply_surface_tension(all, all);
        particles.apply_viscosity(all);
        particles.compute_vorticity(all);
        particles.apply_vorticity(all);
        particles.apply_viscosity(all);
        particles.update_positions(all);
      }
      const bool advanced = !o.graph || frame % 2 == 1 || frame + 1 == o.steps;
      if (!o.write.empty() && advanced && (frame + 1) % o.every == 0) particles.write_positions(o.write);
    }
    particles.sync();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const pbf_stats st = particles.stats();
    std::printf("{\"particles\": %zu, \"steps\": %d, \"mode\": \"%s\", \"seconds\": %.6f, "
                "\"particle_substeps_per_s\": %.6e, \"max_density_error\": %.6f, \"neighbor_overflow\": %llu}\n",
                n, o.steps, o.tolerance >= 0.0f ? "solve" : o.graph ? "graph" : o.fused ? "fused" : "per-pass", secs,
                (double)n * o.steps / secs,
                st.max_density_error, (unsigned long long)st.neighbor_overflow);
    return 0;
  } catch (const std::exception& e) {
    std::printf("Aborting: %s\n", e.what());  // Simulation/Source/main.cpp:139-141
    return 1;
  }
}
