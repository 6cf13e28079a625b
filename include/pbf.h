/* pbf.h -- C ABI of the MI355X position-based-fluids library.
 *
 * This header is the drop-in boundary for the PBF step of AdamSimpson/SPH-CPP.
 * The reference drives its step through header-only templates:
 *   Common/sim_algorithms.h:43-258   for_each_index, sort_by_key,
 *                                    lower_bound, upper_bound (Thrust)
 *   Simulation/Source/particles.h    one for_each_index lambda per pass
 *   Simulation/Source/neighbors.h    bins, stable sort, bounds, capped lists
 * Device lambdas cannot cross an extern "C" boundary, so each lambda body
 * becomes one entry point taking an index span, the Thrust primitives become
 * concretely typed device-pointer entry points, and the sim::Array /
 * ManagedAllocation storage (Common/array.h, Common/managed_allocation.h)
 * becomes device buffers owned by a context.
 *
 * Conventions
 *   - Every function returns PBF_OK (0) or an error code; pbf_last_error()
 *     gives the message for the calling thread.  Nothing throws.
 *   - Particle indices are the reference's indices: the order particles were
 *     added with pbf_add (particles.h:184-199).  The library stores particles
 *     sorted by cell internally; every getter returns original order.
 *   - Vectors are 3 floats per particle (Vec<float,3>, 12 bytes, vec.h:184-256).
 *   - Work is stream ordered on the context's stream; getters synchronise.
 *     The reference synchronises after every call (sim_algorithms.h:49,60,79,98,110);
 *     here only pbf_sync and the getters do.
 *   - One context per host thread.
 */
#ifndef PBF_H
#define PBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pbf_status {
  PBF_OK = 0,
  PBF_ERR_ARG = 1,        /* bad argument (span out of range, null pointer) */
  PBF_ERR_CAPACITY = 2,   /* "Not enough capacity to push_back" (array.h:127-158) */
  PBF_ERR_HIP = 3,        /* HIP runtime error */
  PBF_ERR_STATE = 4,      /* call order (e.g. a neighbour pass before find_neighbors) */
  PBF_ERR_UNSUPPORTED = 5 /* configuration outside what the library handles */
};

/* IndexSpan = Vec<size_t,2>{begin, end} (vec.h:706), half open. */
typedef struct pbf_index_span {
  uint64_t begin;
  uint64_t end;
} pbf_index_span;

/* Parameters<float,3> public fields used by the step (parameters.h:550-573). */
typedef struct pbf_params {
  uint64_t max_particles_local;  /* capacity (parameters.h:550) */
  uint32_t solve_step_count;     /* Jacobi iterations per step (parameters.h:552) */
  uint32_t max_neighbors;        /* list cap, MAX_NEIGHBORS = 60 (neighbors.h:39); 1..128 */
  float particle_rest_spacing;
  float particle_radius;
  float smoothing_radius;        /* h */
  float neighbor_bin_spacing;    /* cell size and neighbour radius (neighbors.h:198) */
  float rest_density;
  float rest_mass;
  float gravity;                 /* applied to y (particles.h:268) */
  float gamma;                   /* surface tension; 0 skips the pass */
  float lambda_epsilon;
  float k_stiff;                 /* parsed by the reference, never used */
  float visc_c;
  float time_step;
  float max_speed;
  float vorticity_coef;
  float boundary_min[3];         /* must be the origin (neighbors.h:110-113) */
  float boundary_max[3];
  float mover_center[3];
  uint32_t options;              /* PBF_OPT_* bits (no counterpart in the reference) */
} pbf_params;

/* Count neighbour candidates dropped by the list cap (pbf_stats.neighbor_overflow).
 * The reference truncates silently (neighbors.h:223); counting needs the full
 * 27-cell walk, so without this option a full list ends the walk early.  Lists
 * are identical either way. */
#define PBF_OPT_COUNT_OVERFLOW 1u
/* pbf_solve compares the tolerance with the mean density error
 * (1/n) sum_p max(rho_p/rho0 - 1, 0) instead of the maximum. */
#define PBF_OPT_SOLVE_MEAN_ERROR 2u
/* The first density pass of a warm-started pbf_solve also stores the learned warm
 * start's features in device memory (pbf_get_warm_features,
 * pbf_warm_model_add_from_context).  Without it they stay in registers: the
 * learned warm start evaluates its network inside that pass. */
#define PBF_OPT_KEEP_WARM_FEATURES 4u
/* The fused step (pbf_step, pbf_step_graph, pbf_solve, pbf_solve_graph) takes the
 * hardware square root and reciprocal (within about 1 ulp) in the neighbour sums of its
 * pressure solve (densities, multipliers, position updates) in place of the correctly
 * rounded square root and divide; XSPH and vorticity stay exact.  Cell keys, particle order
 * and neighbour lists stay bit-exact; positions then match the exact mode within the
 * tolerance tests/test_gpu_fast_math.py states.  The per-pass entry points stay exact. */
#define PBF_OPT_FAST_MATH 8u
/* A warm-started pbf_solve falls back to a zero warm start when either density error
 * measure grows over the warm update (the max and the mean), not only the one the
 * tolerance applies to. */
#define PBF_OPT_WARM_STRICT 16u
/* The fused step walks every neighbour list in full.  By default, after building the
 * lists it also keeps their entries within 1.1 h (the skin lists, in list order) and its
 * Jacobi passes walk those while no particle has moved 0.045 h since the lists were built:
 * an entry outside them is then farther than h and adds nothing, so the sums are the same,
 * bit for bit.  This option turns the skin lists off (for comparison). */
#define PBF_OPT_FULL_LISTS 32u
/* The list builder walks every particle's 27 cells with one lane of its own.  By default
 * the particles of ordinary cells are listed by a walk shared by all particles of a cell
 * (each candidate tested by one lane against the cell's particles in turn); the lists are
 * the same, bit for bit.  This option keeps the per-lane walk for every particle (for
 * comparison). */
#define PBF_OPT_LANE_LISTS 64u
/* pbf_step and pbf_solve run their Jacobi iterations (the density, dirty-marking,
 * lambda-repair and pressure-update passes of every iteration or solve slot) in one
 * cooperative launch with grid-wide barriers between the passes, instead of one launch per
 * pass; a solve stops on the device.  Same results, bit for bit.  Used for up to 256
 * particles per compute unit (65,536 on a 256-CU device: one block of the launch per CU);
 * larger contexts, graphs being captured (pbf_step_graph, pbf_solve_graph) and solves with
 * PBF_OPT_WARM_RETRY keep the launches per pass (pbf_stats.persistent_grid tells which ran).  A device without
 * cooperative launches makes pbf_step and pbf_solve fail with PBF_ERR_UNSUPPORTED. */
#define PBF_OPT_PERSISTENT 128u
/* A warm-started pbf_solve whose warm start grows the error at slot 0 retries it once with the
 * warm multipliers halved (from the same position stars) before it falls back to a zero warm
 * start: slot 1 checks the damped start against the same error (pbf_solve_info's
 * warm_start_kept is then 2, and the halved multipliers join the sums).  Without it the solve
 * falls back to zero at once. */
#define PBF_OPT_WARM_RETRY 256u

typedef struct pbf_context pbf_context;

/* Most iteration slots of one pbf_solve call. */
#define PBF_MAX_SOLVE_ITERATIONS 256

/* Counters for the last step / pass. */
typedef struct pbf_stats {
  uint64_t neighbor_overflow;  /* candidates within the bin spacing dropped by the cap
                                  (the reference truncates silently, neighbors.h:223) */
  uint64_t outside_grid;       /* particles whose cell id is not in the grid */
  uint64_t density_hacks;      /* triggers of the near-coincidence fix (particles.h:303-305) */
  uint64_t iterations;         /* Jacobi iterations run in the last step */
  double max_density_error;    /* max_p max(rho_p/rho0 - 1, 0) of the last density pass */
  uint64_t tile_listed;        /* particles the last list build listed with the cell-tile walk
                                  (PBF_OPT_LANE_LISTS); counted only with PBF_OPT_COUNT_OVERFLOW */
  uint64_t persistent_grid;    /* blocks of the last cooperative Jacobi launch (PBF_OPT_PERSISTENT; 0: none) */
  uint64_t persistent_capacity; /* blocks of that kernel the device runs at once (0 without the option) */
} pbf_stats;

/* Thread-local message for the last failing call. */
const char* pbf_last_error(void);

/* Fill *p with the INI-less defaults, then derive like
 * Parameters::derive_from_input (parameters.h:123-148): particle_radius,
 * smoothing_radius (1.8 spacing if <= 0), neighbor_bin_spacing (1.2 h if <= 0),
 * rest_mass = 1, rest_density = 1/spacing^3, max_speed = 0.5 h I / dt. */
int pbf_params_derive(pbf_params* p);

/* Particles(parameters) (particles.h:47-57) + Neighbors(parameters)
 * (neighbors.h:67-76): allocates capacity max_particles_local on the device
 * (below 2^28: the neighbour passes address particle arrays with 32-bit byte
 * offsets; PBF_ERR_ARG otherwise). */
int pbf_create(const pbf_params* params, pbf_context** out);
int pbf_destroy(pbf_context* ctx);

/* Replace the parameters (the renderer re-broadcasts them,
 * Simulation/Source/main.cpp:62-63).  Capacity may not change and max_neighbors may not
 * exceed the value the context was created with; a new boundary or bin spacing re-sizes
 * the grid (in a batched context, every scene's grid). */
int pbf_set_params(pbf_context* ctx, const pbf_params* params);
int pbf_get_params(const pbf_context* ctx, pbf_params* out);

/* Particles::add(positions, position_stars, velocities, count)
 * (particles.h:184-199); densities, lambdas and scratch start at 0. */
int pbf_add(pbf_context* ctx, const float* positions, const float* position_stars,
            const float* velocities, uint64_t count);
/* Particles::remove(count) (particles.h:146-154): drop the last count particles. */
int pbf_remove(pbf_context* ctx, uint64_t count);
uint64_t pbf_local_count(const pbf_context* ctx);   /* particles.h:76 */
uint64_t pbf_available(const pbf_context* ctx);     /* particles.h:81 */
uint64_t pbf_max_local_count(const pbf_context* ctx);

/* Array getters (particles.h:91-141) in original order; 3 floats per
 * particle for vectors, 1 for scalars.  Setters overwrite all local particles. */
int pbf_get_positions(pbf_context* ctx, float* out);
int pbf_get_position_stars(pbf_context* ctx, float* out);
int pbf_get_velocities(pbf_context* ctx, float* out);
int pbf_get_densities(pbf_context* ctx, float* out);
int pbf_get_lambdas(pbf_context* ctx, float* out);
int pbf_get_scratch(pbf_context* ctx, float* out);
int pbf_set_positions(pbf_context* ctx, const float* in);
int pbf_set_position_stars(pbf_context* ctx, const float* in);
int pbf_set_velocities(pbf_context* ctx, const float* in);
int pbf_set_densities(pbf_context* ctx, const float* in);
int pbf_set_lambdas(pbf_context* ctx, const float* in);
int pbf_set_scratch(pbf_context* ctx, const float* in);

/* ---- one entry per for_each_index body (spans in original indices) ---- */
int pbf_apply_external_forces(pbf_context* ctx, pbf_index_span span);   /* particles.h:263-270 */
int pbf_predict_positions(pbf_context* ctx, pbf_index_span span);       /* particles.h:275-287 */
/* Particles::find_neighbors -> Neighbors::find (particles.h:256-258,
 * neighbors.h:241-248): calculate_bins, stable sort_by_key, lower/upper
 * bound, fill_neighbors.  bin_span.begin must be 0 (neighbors.h:130-132). */
int pbf_find_neighbors(pbf_context* ctx, pbf_index_span bin_span, pbf_index_span fill_span);
int pbf_compute_densities(pbf_context* ctx, pbf_index_span span);       /* particles.h:292-313 */
int pbf_compute_pressure_lambdas(pbf_context* ctx, pbf_index_span span);/* particles.h:318-342 */
int pbf_compute_pressure_dps(pbf_context* ctx, pbf_index_span span, int substep); /* :348-358 */
int pbf_update_position_stars(pbf_context* ctx, pbf_index_span span);  /* particles.h:363-370 */
int pbf_update_velocities(pbf_context* ctx, pbf_index_span span);       /* particles.h:375-385 */
int pbf_apply_surface_tension(pbf_context* ctx, pbf_index_span color_field_span,
                              pbf_index_span surface_tension_span);     /* particles.h:400-431 */
int pbf_apply_viscosity(pbf_context* ctx, pbf_index_span span);         /* particles.h:436-447 */
int pbf_compute_vorticity(pbf_context* ctx, pbf_index_span span);       /* particles.h:452-466 */
int pbf_apply_vorticity(pbf_context* ctx, pbf_index_span span);         /* particles.h:471-487 */
int pbf_update_positions(pbf_context* ctx, pbf_index_span span);        /* particles.h:390-394 */

/* The whole step body of Simulation/Source/main.cpp:71-124 on one domain
 * (resident = local = all particles), nsteps times, with fused kernels. */
int pbf_step(pbf_context* ctx, uint32_t nsteps);

/* ---- batched scenes (no counterpart in the reference) ----
 * Independent scenes in one context and one set of launches.  Every parameter is shared
 * except the box: scene s occupies [0, boxes[3s .. 3s+2]] in its own coordinates, with its
 * own grid derived as for a single scene (neighbors.h:69-71).  A particle's cell key is its
 * scene's first key plus the key in its scene's grid, so the sort keeps scenes apart and a
 * neighbour walk never leaves its scene: every scene evolves exactly as it would alone
 * (the max density error and the tolerance of pbf_solve are taken over all scenes).
 * pbf_add adds to scene 0; pbf_grid_dims reports scene 0's grid; pbf_num_cells the total. */
int pbf_create_batched(const pbf_params* params, uint32_t num_scenes, const float* boxes,
                       pbf_context** out);
int pbf_add_to_scene(pbf_context* ctx, uint32_t scene, const float* positions,
                     const float* position_stars, const float* velocities, uint64_t count);
uint32_t pbf_num_scenes(const pbf_context* ctx);

/* ---- iterate-to-tolerance solve (no counterpart in the reference, which runs a
 * fixed solve_step_count of Jacobi iterations, Simulation/Source/main.cpp:82) ----
 * One step like pbf_step, except the pressure loop: iteration slot k runs the density
 * pass and measures e_k = max_p max(rho_p/rho0 - 1, 0) (the clamp of particles.h:325);
 * if e_k <= tolerance the loop ends (x* keeps that pass's near-coincidence fix),
 * otherwise lambdas, dps and position stars follow.  At most max_iterations slots.
 * Velocities then come from update_velocities (particles.h:375-385).
 *
 * Warm start: before slot 0, a density pass measures e_pre and one dp pass with the
 * warm multipliers Lambda_p in place of lambda moves x*.  If slot 0 then measures
 * e_0 > e_pre, x* returns to its value before the warm update and slot 0 applies
 * nothing (the fallback to a zero warm start).  Sources of Lambda:
 *   PBF_WARM_GIVEN     pbf_set_warm_lambdas (original particle order)
 *   PBF_WARM_PREVIOUS  the multipliers summed over the previous solve or step
 * Every solve keeps the sum of the multipliers it applied (the warm ones included);
 * pbf_get_lambda_sums returns it.  Slots are queued in chunks of 8; between chunks
 * the call waits for the device to say whether the solve has ended (pbf_solve_graph
 * queues every slot instead and never waits). */
enum { PBF_WARM_NONE = 0, PBF_WARM_GIVEN = 1, PBF_WARM_PREVIOUS = 2, PBF_WARM_LEARNED = 3 };

/* The error e_k compared with the tolerance is the maximum by default and the mean with
 * PBF_OPT_SOLVE_MEAN_ERROR; both are recorded.  The mean is summed in fixed point
 * (each particle's error times 2^32, truncated to an integer), so it does not depend
 * on the order of the sum. */
typedef struct pbf_solve_info {
  uint32_t iterations;              /* slots used (<= max_iterations) */
  int32_t warm_start_kept;          /* 1: warm-start positions kept; 2: its damped retry's kept
                                       (PBF_OPT_WARM_RETRY); 0: rejected or none */
  float max_error_before_warm_start;
  float mean_error_before_warm_start;
  float max_error_history[PBF_MAX_SOLVE_ITERATIONS];   /* of every used slot */
  float mean_error_history[PBF_MAX_SOLVE_ITERATIONS];
} pbf_solve_info;

int pbf_solve(pbf_context* ctx, float tolerance, uint32_t max_iterations, int warm_start);
int pbf_set_warm_lambdas(pbf_context* ctx, const float* lambdas);
int pbf_get_lambda_sums(pbf_context* ctx, float* out);
/* replace the kept sums (original order): the next solve's PBF_WARM_PREVIOUS source
 * and feature 2 of the learned warm start */
int pbf_set_lambda_sums(pbf_context* ctx, const float* in);
int pbf_get_solve_info(pbf_context* ctx, pbf_solve_info* out);

/* No counterpart in the reference (the loop of Simulation/Source/main.cpp:82-99 run
 * to a tolerance, on the device).  nsteps pbf_solve steps with no host round trip:
 * two steps are captured into one HIP graph and replayed (an odd count ends with one
 * step queued the same way).  Every slot of a step is queued; once a slot's density
 * error is <= tolerance the later slots' kernels return at once, so the call never
 * waits for the device.  Results are bit for bit those of pbf_solve.  The graph is
 * captured again when the particle count, parameters, tuning, tolerance,
 * max_iterations, warm_start or warm-start gain change. */
int pbf_solve_graph(pbf_context* ctx, float tolerance, uint32_t max_iterations, int warm_start,
                    uint32_t nsteps);

/* One record per pbf_solve / pbf_solve_graph step, written by the device at the
 * end of the step (so graph replays record every step).  The context keeps the
 * last PBF_SOLVE_LOG_SIZE. */
typedef struct pbf_solve_record {
  uint64_t index;                   /* 0, 1, 2, ... over the context's solves */
  uint32_t iterations;              /* slots used */
  int32_t warm_start_kept;
  float max_error;                  /* of the last used slot */
  float mean_error;
  float max_error_before_warm_start;
  float mean_error_before_warm_start;
} pbf_solve_record;
#define PBF_SOLVE_LOG_SIZE 4096
/* The last min(count, solves so far, PBF_SOLVE_LOG_SIZE) records, oldest first;
 * *returned = how many were written. */
int pbf_get_solve_log(pbf_context* ctx, pbf_solve_record* out, uint32_t count, uint32_t* returned);

/* ---- learned warm start (PBF_WARM_LEARNED; no counterpart in the reference) ----
 * Features of particle p, from the first density pass of a warm-started solve
 * (x* after external forces and prediction):
 *   0  rho_p/rho0 - 1                      4  neighbours within h / 32
 *   1  lambda_p of that pass / S           5  v_p.y * dt / h
 *   2  previous solve's multiplier sum / S 6  mean (y_q - y_p) over neighbours within h / h
 *   3  sum of |grad C|^2 terms * 1e-5      7  m dt sum_q (v_q - v_p) . grad_q C  (velocity
 *                                             divergence: the density change over one step)
 * with S = PBF_WARM_LAMBDA_SCALE and grad_q C = -(1/rho0) grad W(x_p - x_q) (particles.h:332),
 * limited to [-1,4], [-64,0], [-64,0], [0,8], [0,4], [-4,4], [-2,2], [-4,4].
 * A network 8 -> 64 -> 64 -> 1 (SiLU after both hidden layers; BF16 MFMA operands, FP32
 * accumulation) gives y_p, and the warm multiplier is Lambda_p = Lambda_prev,p + S * y_p:
 * the previous solve's sum, corrected (a network that outputs 0 is PBF_WARM_PREVIOUS).
 * Parameters (PBF_WARM_PARAMS floats): W1[64][8], b1[64], W2[64][64] (row = output
 * unit), b2[64], w3[64], b3.  Training minimises the mean squared error against targets
 * (Lambda_p - Lambda_prev,p) / S, Lambda being the multiplier sum a solve applied
 * (pbf_warm_model_add_from_context), with Adam (beta 0.9, 0.999, eps 1e-8) on FP32 master
 * weights. */
#define PBF_WARM_FEATURES 8
#define PBF_WARM_HIDDEN 64
#define PBF_WARM_PARAMS 4801
#define PBF_WARM_LAMBDA_SCALE 1e-4f

typedef struct pbf_warm_model pbf_warm_model;
int pbf_warm_model_create(uint64_t max_samples, pbf_warm_model** out);
int pbf_warm_model_destroy(pbf_warm_model* model);
int pbf_warm_model_init(pbf_warm_model* model, uint64_t seed);
int pbf_warm_model_set_params(pbf_warm_model* model, const float* params);
int pbf_warm_model_get_params(pbf_warm_model* model, float* params);
/* sample store: features (count x 8) and targets (count, already divided by S) */
int pbf_warm_model_add_samples(pbf_warm_model* model, const float* features, const float* targets,
                               uint64_t count);
/* samples from the context's last warm-started solve, for every stride-th particle in
 * the context's cell order (stride >= 1): its features, and as target the multiplier sum
 * it applied minus the previous sums, over S, limited to [-64, 64] */
int pbf_warm_model_add_from_context(pbf_warm_model* model, pbf_context* ctx, uint32_t stride);
int pbf_warm_model_clear_samples(pbf_warm_model* model);
uint64_t pbf_warm_model_sample_count(const pbf_warm_model* model);
/* gradient and loss of samples [first, first + count) (no update) */
int pbf_warm_model_gradient(pbf_warm_model* model, uint64_t first, uint32_t count, float* grad, float* loss);
/* one Adam update with a given gradient */
int pbf_warm_model_adam_step(pbf_warm_model* model, const float* grad, float learning_rate);
/* epochs of shuffled minibatches; *mean_loss = mean batch loss over the run */
int pbf_warm_model_train(pbf_warm_model* model, uint32_t epochs, uint32_t batch, float learning_rate,
                         uint64_t seed, float* mean_loss);
/* network output (Lambda / S) for count feature rows */
int pbf_warm_model_predict(pbf_warm_model* model, const float* features, uint64_t count, float* out);
/* copy the model's weights into the context (used by PBF_WARM_LEARNED) */
int pbf_set_warm_model(pbf_context* ctx, pbf_warm_model* model);
/* scale of the network's correction: Lambda = Lambda_prev + gain * S * y (default 1, range [0, 4];
 * 0 is PBF_WARM_PREVIOUS) */
int pbf_set_warm_gain(pbf_context* ctx, float gain);
/* features of the last warm-started solve (run with PBF_OPT_KEEP_WARM_FEATURES), n x 8,
 * original order */
int pbf_get_warm_features(pbf_context* ctx, float* out);

/* pbf_step with the launches of two steps captured once into a HIP graph and replayed
 * (an odd count ends with one ordinary step).  Same results as pbf_step.  The graph is
 * captured again when the particle count, the parameters or the buffer roles changed
 * (add, remove, set_params, per-pass calls). */
int pbf_step_graph(pbf_context* ctx, uint32_t nsteps);

/* ---- launch tuning (no counterpart in the reference) ----
 * Launch parameters that do not change any result: the block size of the fused
 * per-particle kernels (64, 128 or 256), the candidate budget after which the neighbour
 * walk of one particle is handed from one lane to a whole wavefront (64..65536, used up to
 * 2040; it applies to particles whose 27 cells hold more than twice that many candidates,
 * or more than 2048), and the number
 * of blocks of that wavefront builder (1..65536).  The bin spacing and the list cap
 * are not tuned: both are part of the reference's neighbour semantics (neighbors.h:39,198).
 * pbf_autotune times `steps` fused steps from the current state for each candidate,
 * restores the state (the following steps are bit-identical to an untuned run), keeps
 * the fastest setting and returns it with its mean step time. */
typedef struct pbf_tuning {
  uint32_t block_size;
  uint32_t heavy_budget;
  uint32_t heavy_blocks;
} pbf_tuning;
int pbf_set_tuning(pbf_context* ctx, const pbf_tuning* tuning);
int pbf_get_tuning(const pbf_context* ctx, pbf_tuning* out);

/* No counterpart in the reference.  Picks the bin spacing (cell size and list radius)
 * and the list cap for speed, for callers that do not need the reference's list layout
 * (bin spacing 1.2 h, neighbors.h:198, cap 60, neighbors.h:39).  Candidate spacings are
 * h, 1.1 h and 1.2 h.  For each, one list build from the current state with the largest
 * cap the context holds gives the longest list; a spacing whose lists do not fit is
 * dropped, and the others get a cap of 1.25x that longest list (rows of 8, at most the
 * created cap).  Every candidate therefore lists every neighbour within h, and every
 * neighbour sum runs over the same particles as without a cap (entries between h and the
 * spacing add exactly zero): layouts differ only in the order of the sums.  Each is timed
 * over `steps` fused steps from the current state, the state is restored, and the fastest
 * is applied with pbf_set_params (read it back with pbf_get_params).  A cap that later
 * overflows is counted with PBF_OPT_COUNT_OVERFLOW.  The reference's own layout
 * truncates lists under compression, so a tuned layout changes results against it:
 * leave the layout alone where results must match the reference's. */
typedef struct pbf_layout {
  float neighbor_bin_spacing;
  uint32_t max_neighbors;
  double ms_per_step;      /* of the chosen layout */
  uint32_t longest_list;   /* at the chosen spacing, on the state it was tuned on */
} pbf_layout;
int pbf_autotune_layout(pbf_context* ctx, uint32_t steps, pbf_layout* best);
int pbf_autotune(pbf_context* ctx, uint32_t steps, pbf_tuning* best, double* best_ms_per_step);

/* Block until the context's queued work is done. */
int pbf_sync(pbf_context* ctx);

int pbf_get_stats(pbf_context* ctx, pbf_stats* out);

/* ---- introspection of the neighbour structures (neighbors.h:279-284) ----
 * Valid after pbf_find_neighbors or pbf_step. */
int pbf_grid_dims(const pbf_context* ctx, uint64_t dims[3]);     /* neighbors.h:253-255 */
uint64_t pbf_num_cells(const pbf_context* ctx);
/* bin_ids_ after sort_bins: sorted cell keys (n entries) */
int pbf_get_bin_ids(pbf_context* ctx, uint64_t* out);
/* particle_ids_ after sort_bins: original index of each sorted slot */
int pbf_get_particle_ids(pbf_context* ctx, uint64_t* out);
/* begin_indices_ / end_indices_ (num_cells entries each) */
int pbf_get_bin_bounds(pbf_context* ctx, uint64_t* begin, uint64_t* end);
/* neighbor_lists_: out[p*max_neighbors + k] = original index of the k-th
 * neighbour of particle p; counts[p] = list length. */
int pbf_get_neighbor_lists(pbf_context* ctx, uint64_t* out, int32_t* counts);

/* ---- sim::algorithms primitives on device memory (sim_algorithms.h) ---- */
/* sort_by_key (sim_algorithms.h:57-61,134-137,207-210): stable LSD radix sort
 * of n (key, value) pairs in place.  key_bits bounds the key width (1..32). */
int pbf_sort_by_key_u32(uint32_t* d_keys, uint32_t* d_values, uint64_t n, uint32_t key_bits);
/* No counterpart in the reference.  pbf_sort_by_key_u32 for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same stable sort and result,
 * with its work space in a fixed device array of the library (about 10 MiB
 * per GPU, held until the process ends) instead of hipMallocAsync.  n above
 * PBF_NOALLOC_MAX_PAIRS returns PBF_ERR_CAPACITY.  Every call on a GPU
 * shares that GPU's work space: calls that overlap on one GPU race. */
#define PBF_NOALLOC_MAX_PAIRS (1ull << 20)
int pbf_sort_by_key_u32_noalloc(uint32_t* d_keys, uint32_t* d_values, uint64_t n, uint32_t key_bits);
/* lower_bound / upper_bound of the counting range [search.begin, search.end)
 * over sorted keys (sim_algorithms.h:69-99,145-173,218-246); one result per
 * searched value. */
int pbf_lower_bound_counting(const uint32_t* d_sorted_keys, uint64_t n, pbf_index_span search,
                             uint32_t* d_result);
int pbf_upper_bound_counting(const uint32_t* d_sorted_keys, uint64_t n, pbf_index_span search,
                             uint32_t* d_result);
/* Device memory for the primitives. */
int pbf_device_alloc(uint64_t bytes, void** out);
int pbf_device_free(void* ptr);
int pbf_memcpy_to_device(void* dst, const void* src, uint64_t bytes);
int pbf_memcpy_to_host(void* dst, const void* src, uint64_t bytes);
int pbf_memcpy_device(void* dst, const void* src, uint64_t bytes);  /* device to device */

/* ---- self-check (no counterpart in the reference) ----
 * The neighbour passes divide and take square roots with the library's own
 * sequences, built to give the bits of the correctly rounded operations for
 * every input.  Compares them with the plain operations on `samples`
 * pseudo-random operand sets (raw bit patterns, operands shaped like the
 * passes' own, log-uniform magnitudes over 2^-48..2^48) and writes out[0] =
 * divide mismatches, out[1] = square-root mismatches, out[2] / out[3] = samples
 * that took the shared-reciprocal divide / the unwrapped square root, out[4] =
 * divide mismatches among the operands shaped like the passes' own. */
int pbf_check_arithmetic(uint64_t samples, uint64_t seed, uint64_t* out);

/* ---- output ----
 * Append one positions record to `path`, with the content AdiosWriter::write_particles
 * writes per call (adios_writer.h:78-114, adios_writer.xml:4-14): int64 global_bytes,
 * int64 local_bytes, int64 offset_bytes, then local_bytes of positions (Vec<float,3>, original
 * order).  ADIOS itself is not used; the file is these records back to back (one domain:
 * global = local, offset = 0). */
int pbf_append_positions_record(pbf_context* ctx, const char* path);

/* ---- measurement ----
 * Bracket every launch of the kernels named in `names` (comma separated,
 * "" = none) with HIP events on the context stream; pbf_kernel_time returns
 * the accumulated launches and milliseconds of one kernel since enabling. */
int pbf_profile_kernels(pbf_context* ctx, const char* names);
int pbf_kernel_time(pbf_context* ctx, const char* name, uint64_t* launches, double* total_ms);
/* Events around a whole region on the context stream. */
int pbf_region_begin(pbf_context* ctx);
int pbf_region_end(pbf_context* ctx, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* PBF_H */
