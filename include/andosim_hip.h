/*
 * andosim_hip.h — C ABI of the MI355X (gfx950) per-step cloth solve.
 *
 * This is the boundary beneath the Python module `ando_barrier_core`: every
 * entry point takes plain pointers and sizes, owns its device memory inside an
 * opaque context, and returns 0 on success or a negative ANDO_E* code with a
 * message readable through ando_last_error().  Solver non-convergence is not
 * an error (the reference only prints it); it is reported in ando_stats.
 * Calls on one context are serialised (each holds the context's lock for the
 * whole call); distinct contexts may be used from different threads at once.
 *
 * Reference interfaces replaced (paths relative to the AndoSim source tree):
 *   ando_set_mesh          Mesh::initialize + compute_rest_state       src/core/mesh.cpp:8-75
 *   ando_set_mesh_vertices Mesh::set_positions / Mesh.vertices setter   src/core/mesh.cpp:122-124,
 *                                                                       src/py/bindings.cpp:166-192
 *   ando_lumped_masses     State::compute_lumped_masses                 src/core/state.cpp:19-33
 *   ando_set_state         State::initialize / State.set_velocities    src/core/state.cpp:7-17,
 *                                                                       src/py/bindings.cpp:268-275
 *   ando_get_state         State.get_positions / get_velocities        src/py/bindings.cpp:248-267
 *   ando_set_previous_positions  State::positions_prev                  src/core/state.h:18, state.cpp:13
 *   ando_set_pins          Constraints::add_pin                         src/core/constraints.cpp:5-11
 *   ando_set_walls         Constraints::add_wall                        src/core/constraints.cpp:13-20
 *   ando_apply_gravity     State.apply_gravity                          src/py/bindings.cpp:284-300
 *   ando_step              Integrator::step                             src/core/integrator.cpp:70-150,
 *                                                                       src/py/bindings.cpp:469-494
 *   ando_compute_contacts  Integrator::compute_contacts                 src/core/integrator.cpp:238-244,
 *                                                                       src/py/bindings.cpp:495-509
 *   ando_energy_diagnostics  EnergyTracker::compute                   src/core/energy_tracker.cpp:7-76
 *   ando_contact_metrics   CollisionValidator::compute_metrics          src/core/collision_validator.cpp:133-157
 *   ando_cfl_inputs        AdaptiveTimestep::compute_max_velocity +
 *                          compute_min_edge_length                      src/core/adaptive_timestep.cpp:82-150
 *   ando_host_alloc/_free  (new, additive) pinned host buffers: ando_get_state into one is a
 *                          direct DMA; the module returns get_positions / get_velocities
 *                          as arrays over such buffers (the reference copies out of an
 *                          Eigen vector, src/py/bindings.cpp:248-267)
 *   ando_step_batch        (new, additive) ando_step on many contexts at once; the reference
 *                          steps independent cloths one after another
 *   ando_get_stats         (new, additive) solver statistics; the reference exposes none
 *                          (src/core/pcg_solver.h:26-27 returns bool only)
 *   ando_linearize         one Newton linearisation: Integrator::compute_gradient +
 *                          assemble_system_matrix                      src/core/integrator.cpp:246-544
 *   ando_pcg               PCGSolver::solve on a caller-supplied BSR   src/core/pcg_solver.cpp:8-146
 *   ando_apply_restitution Integrator::apply_contact_restitution       src/core/integrator.cpp:565-674
 *   ando_get_pattern       block sparsity of the system matrix (the CSR pattern that
 *                          setFromTriplets builds at integrator.cpp:539)
 */
#ifndef ANDOSIM_HIP_H
#define ANDOSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANDO_OK 0
#define ANDO_EINVAL -1   /* bad argument (the reference raises ValueError) */
#define ANDO_ESTATE -2   /* call order violated (the reference raises RuntimeError) */
#define ANDO_EDEVICE -3  /* HIP runtime failure */

typedef struct ando_ctx ando_ctx;

/* Material (src/core/types.h:51-57). */
typedef struct {
  float youngs_modulus;
  float poisson_ratio;
  float density;
  float thickness;
  float bending_stiffness;
} ando_material;

/* SimParams (src/core/types.h:60-100), every field including the ones the
 * Python binding does not expose. */
typedef struct {
  float dt;
  float beta_max;
  int32_t min_newton_steps;
  int32_t max_newton_steps;
  int32_t friction_min_newton_steps;
  float pcg_tol;
  int32_t pcg_max_iters;
  float contact_gap_max;
  float wall_gap;
  int32_t enable_ccd;
  float contact_normal_epsilon;
  float barrier_tolerance;
  int32_t enable_friction;
  float friction_mu;
  float friction_epsilon;
  float friction_tangent_threshold;
  float velocity_damping;
  float contact_restitution;
  int32_t enable_strain_limiting;
  float strain_limit;
  float strain_tau;
  float strain_epsilon;
  float strain_svd_epsilon;
  float hessian_epsilon;
  float min_gap;
} ando_params;

/* Rigid body in/out of a step (src/core/rigid_body.h:16-86).  The library
 * reads the pose and local geometry, and writes back pose and velocities after
 * Integrator::apply_rigid_coupling (integrator.cpp:676-754). */
typedef struct {
  int32_t num_vertices;
  int32_t num_triangles;
  const float* local_vertices; /* (num_vertices, 3), body frame */
  const int32_t* triangles;    /* (num_triangles, 3) */
  float mass;
  float inertia_body_inv[9];   /* row-major */
  float position[3];
  float rotation[9];           /* row-major */
  float linear_velocity[3];
  float angular_velocity[3];
} ando_rigid_body;

/* ContactPair (src/core/collision.h:78-103). */
typedef struct {
  int32_t type; /* 0 point-triangle, 1 edge-edge, 2 wall, 3 rigid point-triangle */
  int32_t idx[4];
  int32_t rigid_body_index;
  int32_t vertex_count;
  float gap;
  float normal[3];
  float witness_p[3];
  float witness_q[3];
  float weights[4];
} ando_contact;

#define ANDO_MAX_RECORDED_SOLVES 64

/* Solver statistics of the last ando_step (new, additive API). */
typedef struct {
  int32_t newton_iterations;
  int32_t pcg_iterations;           /* summed over every PCG solve of the step */
  int32_t pcg_solves;
  int32_t pcg_failures;             /* cap reached or |p.Ap| < 1e-16 */
  int32_t line_search_failures;     /* alpha = 0 */
  int32_t contacts;                 /* contacts detected at the start of the step */
  int32_t rigid_contacts;
  int32_t strain_active;            /* active strain constraints at the last linearisation */
  float beta;
  float last_alpha;
  int32_t pcg_per_solve[ANDO_MAX_RECORDED_SOLVES];
  double step_ms;                   /* wall time of the step, host clock */
  int32_t timed_launches[2];        /* ANDO_MODE_KERNEL_TIMING: launches that ran (not skipped) */
  double timed_ms[2];               /* summed event durations of those launches */
  int32_t timed_iterations;         /* reserved (0) */
  int32_t pcg_kernel;               /* two kernels per iteration with 0: the BSR row SpMV,
                                       3: the sliced-ELL SpMV, 4: the two-level patch kernels
                                       (ANDO_MODE_TWO_LEVEL; 1 and 2 are not used) */
  int32_t learned_solves;           /* solves that used the learned preconditioner */
  int32_t learned_fallbacks;        /* of those, solves that fell back to plain PCG */
  int32_t fallback_iterations;      /* plain PCG iterations spent in fallbacks (included in pcg_iterations) */
  int32_t spmv_lanes;               /* pcg_kernel 3: lanes per row of the sliced-ELL SpMV (1, 2 or 4) */
  int32_t spmv_wg_slices;           /* pcg_kernel 3: slices of 64 rows per SpMV workgroup (0: half a slice) */
  int32_t batched_solves;           /* solves run inside a batched PCG launch (ando_step_batch) */
  int32_t pcg_vec_block;            /* threads per workgroup of the PCG vector kernels */
  int32_t pcg_autotuned;            /* 1: the launch shape above was picked by the autotuner
                                       (environment ANDO_PCG_AUTOTUNE=1, read by ando_create) */
  int32_t long_hit_sorts;           /* detection passes in which a query had more than 64 hits,
                                       sorted by a segmented radix sort instead of by the query */
  int32_t max_candidates;           /* largest candidate list of one detection pass (PT, EE or rigid) */
  int32_t held_iterations;          /* corrected physics: Newton iterations that held the vertices of
                                       pairs within their separation that would not move apart */
  int32_t capped_trials;            /* corrected physics: line searches whose first trial motion was
                                       halved (once per count) so that its detection holds at most 2^27
                                       candidates */
  double timed_coupling_bytes;      /* ANDO_MODE_KERNEL_TIMING: algorithmic bytes of the self-contact
                                       coupling workgroups in the timed SpMV launches */
} ando_stats;

int32_t ando_abi_version(void);
const char* ando_last_error(void);

int ando_create(int32_t device, ando_ctx** out);
int ando_destroy(ando_ctx* ctx);

int ando_set_mesh(ando_ctx* ctx, const float* vertices, int32_t num_vertices, const int32_t* triangles,
                  int32_t num_triangles, const ando_material* material);
int ando_set_mesh_vertices(ando_ctx* ctx, const float* vertices, int32_t num_vertices);
/* Host helper: State::compute_lumped_masses (src/core/state.cpp:19-33) for a
 * mesh, from its rest areas (mesh.cpp:18-75).  Needs no context and no GPU. */
int ando_lumped_masses(const float* vertices, int32_t num_vertices, const int32_t* triangles,
                       int32_t num_triangles, const ando_material* material, float* masses_out);
int ando_set_state(ando_ctx* ctx, const float* positions, const float* velocities, const float* masses,
                   int32_t num_vertices);
int ando_get_state(ando_ctx* ctx, float* positions, float* velocities, int32_t num_vertices);
/* Page-locked host memory (new, additive).  ando_get_state copies into such a
 * buffer by DMA, without the staging copy that pageable memory needs.  Needs no
 * context; ando_host_free(NULL) does nothing. */
int ando_host_alloc(size_t bytes, void** out);
int ando_host_free(void* ptr);
/* State::positions_prev (src/core/state.h:18): set once by State::initialize
 * (state.cpp:13) and read by friction (integrator.cpp:39,384).  Defaults to the
 * positions of the first complete ando_set_state. */
int ando_set_previous_positions(ando_ctx* ctx, const float* positions_prev, int32_t num_vertices);
int ando_set_pins(ando_ctx* ctx, const int32_t* vertices, const float* targets, int32_t count);
int ando_set_walls(ando_ctx* ctx, const float* normals, const float* offsets, const float* gaps, int32_t count);

/* Precision selection (new, additive).  ANDO_MODE_FP64_NEAR_CONTACT (default
 * on) evaluates barrier derivatives in FP64 for gaps below 1e-2 * gap_max;
 * with flags = 0 every barrier term rounds exactly as the reference's FP32
 * build. */
#define ANDO_MODE_FP64_NEAR_CONTACT 1
/* ANDO_MODE_KERNEL_TIMING brackets every PCG launch with HIP events on the
 * context's stream and accumulates their durations into ando_stats.timed_ms
 * (index 0 the SpMV kernel, 1 the update kernel).  Flag value 4 is not used. */
#define ANDO_MODE_KERNEL_TIMING 2
/* ANDO_MODE_LEARNED_PRECOND (needs ando_set_precond_weights) solves each
 * Newton system with the learned block preconditioner and the warm start it
 * predicts; a solve that misses the tolerance is redone with plain
 * block-Jacobi PCG from x0 = 0 (counted in ando_stats.learned_fallbacks). */
#define ANDO_MODE_LEARNED_PRECOND 8
/* With ANDO_MODE_LEARNED_PRECOND, ANDO_MODE_LEARNED_WARM_START_ONLY keeps the
 * block-Jacobi preconditioner and uses the network only for the warm start. */
#define ANDO_MODE_LEARNED_WARM_START_ONLY 16
/* ANDO_MODE_CORRECTED_CONTACT (new; the reference's self broad phase is always
 * empty, collision.cpp:63-160) detects self point-triangle and edge-edge
 * contacts with the predicate the reference intends (point box +-1e-4, edge
 * boxes, gap < 0.01) through a GPU LBVH, and puts both barriers in the energy;
 * ando_compute_contacts then lists them before the rigid contacts. */
#define ANDO_MODE_CORRECTED_CONTACT 32
/* The two-kernel PCG multiplies by the system matrix in a sliced-ELL copy of
 * its values (slices of 64 rows, made once per solve); ANDO_MODE_PCG_ROW_SPMV
 * reads the BSR values in place instead (8 lanes per block row).  Both sum a
 * row's blocks in the same order and give the same product bit for bit. */
#define ANDO_MODE_PCG_ROW_SPMV 64
/* ANDO_MODE_STATIC_RIGIDS (new): the rigid bodies are fixed colliders.  Their
 * contacts still act on the cloth, but Integrator::apply_rigid_coupling
 * (integrator.cpp:676-754) is skipped: no contact forces on the bodies and no
 * RigidBody::integrate, so the bodies passed to ando_step are left unchanged.
 * Meant for obstacles the reference models as very dense bodies. */
#define ANDO_MODE_STATIC_RIGIDS 128
/* ANDO_MODE_ADDITIVE_CCD (new; with ANDO_MODE_CORRECTED_CONTACT): the line
 * search tests every self contact's motion over a trial step with additive
 * CCD (a conservative time of impact, Li et al. 2021) instead of the
 * reference's 10 distance samples (line_search.cpp:182-243), which a fast
 * primitive can pass between.  A trial is feasible only if the whole motion
 * keeps every contact pair apart. */
#define ANDO_MODE_ADDITIVE_CCD 256
/* ANDO_MODE_CORRECTED_PHYSICS (new; implies ANDO_MODE_CORRECTED_CONTACT) steps
 * the physics the reference intends where its step does not:
 *   - membrane force from the current positions (the reference evaluates F
 *     from the rest mesh, elasticity.cpp:18-61, so cloth has no restoring
 *     force) with the Gauss-Newton Hessian 2k (B B^T)_ab (I - n n^T);
 *   - pins attract: a spring of stiffness 1e4 m / dt^2 to the target (the
 *     reference's pin barrier pushes away from it, barrier.cpp:161-215);
 *   - gravity is implicit: ando_apply_gravity updates velocities only and the
 *     step's x_hat = x + dt v carries the motion (the reference also moves the
 *     vertices before the solve, bindings.cpp:284-300, past any contact);
 *   - contact candidates from the vertices' swept boxes, grown by gap_max,
 *     re-detected after every solve; gaps, normals and weights at every Newton
 *     iteration; contact stiffness m / max(gap_max, min_gap)^2 + n^T H n;
 *   - the line search bounds alpha by every candidate's additive-CCD time of
 *     impact, rigid triangles included, and lets no vertex get more violated
 *     at the first wall; v = (x - x_old) / dt; rigid bodies are pushed away
 *     from the cloth.
 * Setting or clearing the flag resets the cached step constants. */
#define ANDO_MODE_CORRECTED_PHYSICS 512
/* ANDO_MODE_TWO_LEVEL (new): the step's Newton systems (and ando_pcg) are
 * solved by CG with a two-level Schwarz preconditioner instead of block-Jacobi
 * (pcg_solver.cpp:8-146): the block rows in Morton order of the rest positions,
 * cut into patches of 256 to 512 rows; per patch a fixed odd-degree Chebyshev
 * polynomial in D^-1 A_pp (block-Jacobi scaling, the patch's own blocks;
 * degree ANDO_TL_CHEB, default 9, over [lmax / ANDO_TL_RATIO, lmax],
 * ratio default 5), plus a coarse correction with three
 * unknowns per patch (A_c = P^T A P, inverted in FP64 before the solve's
 * iterations; for up to 1024 patches).  In ando_step the inverse is rebuilt
 * on every ANDO_TL_COARSE_EVERY-th solve (default 2) and the solves between
 * keep the last one, which leaves the preconditioner symmetric positive
 * definite and fixed within each solve (the count restarts at each
 * ando_set_state, so a step depends only on the state it starts from);
 * ando_pcg always rebuilds it.  In
 * ando_step a Newton solve after the first of an inner Newton loop starts
 * from (1 - 1.25 alpha) times the last direction (alpha: the last accepted
 * line-search step), falling back to x0 = 0 when that start's residual is no
 * smaller than b's (ANDO_TL_WARM=0: always x0 = 0).  Same
 * stopping test (max-norm relative residual, |p.Ap| < 1e-16 breakdown, the
 * iteration cap), far fewer iterations on cloth systems.  ando_stats.pcg_kernel
 * reads 4. */
#define ANDO_MODE_TWO_LEVEL 1024
int ando_set_mode(ando_ctx* ctx, int32_t flags);

int ando_apply_gravity(ando_ctx* ctx, const float* gravity3, float dt);
int ando_step(ando_ctx* ctx, const ando_params* params, ando_rigid_body* bodies, int32_t num_bodies);
int ando_get_stats(ando_ctx* ctx, ando_stats* out);

/* Batched stepping (new; the reference has no batched API and steps
 * independent cloths one after another): ando_step on `count` distinct
 * contexts, each on its own stream, from up to `threads` host threads
 * (0 = the host's hardware concurrency; at most 256), so small cloths overlap
 * on the GPU; when all contexts are on one device, the PCG solves that the
 * threads reach together run as one batched launch pair per iteration.
 * params[k], bodies[k] / num_bodies[k] (both may be NULL: no bodies) belong to
 * ctxs[k].  Each context's result equals a plain ando_step.  Returns the first
 * failing entry's code, with its index in the message; the others still run. */
int ando_step_batch(ando_ctx* const* ctxs, const ando_params* params, ando_rigid_body* const* bodies,
                    const int32_t* num_bodies, int32_t count, int32_t threads);

int ando_compute_contacts(ando_ctx* ctx, const ando_rigid_body* bodies, int32_t num_bodies, ando_contact* out,
                          int32_t capacity, int32_t* count);

/* Intersection validator (new; the reference's CollisionValidator only counts
 * gap < 0 on unsigned distances, collision_validator.cpp:64, so it never sees
 * a penetration).  edge_triangle: (mesh edge, triangle) pairs sharing no
 * vertex where the edge properly crosses the triangle (double-precision
 * orientation tests, through the LBVH); inside_rigid: cloth vertices whose
 * generalized winding number w.r.t. a rigid body's closed surface exceeds 1/2. */
int ando_count_intersections(ando_ctx* ctx, const ando_rigid_body* bodies, int32_t num_bodies, int32_t* edge_triangle,
                             int32_t* inside_rigid);

/* Per-frame diagnostics the add-on reads after each frame
 * (blender_addon/operators.py:909), from the resident state:
 * EnergyTracker::compute, src/core/energy_tracker.cpp:7-76, with the elastic
 * energy of Elasticity::compute_energy, src/core/elasticity.cpp:5-14,115-126
 * (F from the mesh vertices).  Each per-vertex and per-face term is the
 * reference's float expression; the sums are taken in double (the reference
 * accumulates in float). */
typedef struct {
  double kinetic_energy;
  double elastic_energy;
  double linear_momentum[3];
  double angular_momentum[3];
  double max_velocity;
} ando_energy;
int ando_energy_diagnostics(ando_ctx* ctx, ando_energy* out);

/* CollisionValidator::compute_metrics (src/core/collision_validator.cpp:7-157,
 * 178-230; called at blender_addon/operators.py:973) with the resident
 * velocities.  contacts == NULL: over the contacts the last
 * ando_compute_contacts listed, without listing them again (ANDO_ESTATE if a
 * step or another detection has replaced them since).  Otherwise over the
 * caller's `count` contacts.  Counts, minima and maxima are exact; the sums are
 * taken in double and the averages are float quotients of the rounded sums.
 * Rigid contacts (type 3) count in the totals and gap statistics but, as in
 * the reference, carry no relative velocity. */
typedef struct {
  int32_t num_point_triangle;
  int32_t num_edge_edge;
  int32_t num_wall;
  int32_t num_total_contacts;
  int32_t num_penetrations;
  int32_t num_ccd_contacts;
  int32_t num_broad_phase_contacts;
  int32_t num_relative_velocities;
  int32_t has_major_penetration;
  int32_t has_tunneling;
  float min_gap;
  float max_gap;
  float avg_gap;
  float max_penetration;
  float avg_penetration;
  float ccd_effectiveness;
  float max_relative_velocity;
  float avg_relative_velocity;
  double sum_gap;
  double sum_penetration;
  double sum_relative_velocity;
} ando_collision_metrics;
int ando_contact_metrics(ando_ctx* ctx, const ando_contact* contacts, int32_t count, float gap_max,
                         int32_t ccd_enabled, ando_collision_metrics* out);

/* AdaptiveTimestep's two reductions on the resident state, without reading the
 * velocities back (called through AdaptiveTimestep.compute_next_dt at
 * blender_addon/operators.py:872-880): max_velocity as
 * AdaptiveTimestep::compute_max_velocity (src/core/adaptive_timestep.cpp:136-150)
 * over the device velocities, min_edge_length as compute_min_edge_length
 * (:82-134) over the mesh edges and mesh vertices.  Both are exact (a maximum
 * and a minimum of the reference's float expressions). */
int ando_cfl_inputs(ando_ctx* ctx, float* max_velocity, float* min_edge_length);

int ando_get_pattern(ando_ctx* ctx, int32_t* row_ptr, int32_t* cols, int32_t* num_blocks);
int ando_linearize(ando_ctx* ctx, const ando_params* params, const float* x_target, const ando_rigid_body* bodies,
                   int32_t num_bodies, float* gradient, float* values);
int ando_pcg(ando_ctx* ctx, int32_t num_rows, const int32_t* row_ptr, const int32_t* cols, const float* values,
             const float* b, float* x, float tol, int32_t max_iters, int32_t* iterations, int32_t* converged);
/* Allocation-free PCG (new; no counterpart in the reference).  For callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The same solve as ando_pcg
 * in ANDO_MODE_PCG_ROW_SPMV (block-Jacobi PCG, pcg_solver.cpp:8-146; the two
 * return the same x and iteration count bit for bit), on a BSR system already
 * in device memory of `device`: row_ptr (num_rows + 1), cols (num_blocks),
 * values (9 per block, row-major 3x3), b and x (3 * num_rows) are device
 * pointers; x receives the solution from x0 = 0.  iterations / converged
 * (host pointers, nullable) report the outcome.  stream: the hipStream_t to
 * run on (NULL: the null stream); the call waits for it before returning.
 * The library allocates nothing in this call: its work memory is one fixed
 * device array per GPU (about 13 MB, held until the process ends), sized for
 * num_rows <= 131072 and max_iters <= 8192; larger requests fail with
 * ANDO_EINVAL.  Calls on one GPU share that array and are serialised by the
 * library within a process.  The pattern is checked on the device before any
 * kernel follows it (ANDO_EINVAL on a bad row_ptr, a column out of range or a
 * missing diagonal block). */
int ando_pcg_noalloc(int32_t device, int32_t num_rows, int32_t num_blocks, const int32_t* row_ptr, const int32_t* cols,
                     const float* values, const float* b, float* x, float tol, int32_t max_iters, void* stream,
                     int32_t* iterations, int32_t* converged);
/* The restitution pass that ando_step runs after its velocity update when
 * contact_restitution > 0, on the current state: contacts detected at the
 * current positions (self contacts, then rigid), impulses in list order
 * (self, then rigid, which also change the bodies' velocities), then walls. */
int ando_apply_restitution(ando_ctx* ctx, const ando_params* params, ando_rigid_body* bodies, int32_t num_bodies);

/* ---- Learned preconditioner (new; no counterpart in the reference, whose PCG
 * is block-Jacobi from x0 = 0: pcg_solver.cpp:8-75, integrator.cpp:187). ----
 * A mesh message-passing network (4 layers, width 64, BF16 weights, FP32
 * accumulation) maps each Newton system to per-vertex 3x3 preconditioner
 * blocks P_i and the warm start x0 = alpha P b (alpha minimises the A-norm
 * error along P b).  Weights are a flat float array of
 * ando_nn_num_weights() values. */
int32_t ando_nn_num_weights(void);
int ando_set_precond_weights(ando_ctx* ctx, const float* weights, int32_t count);

/* Trainer: holds Newton systems (BSR, b) on one device and trains the network
 * on them with the library's own forward and backward kernels and Adam. */
typedef struct ando_trainer ando_trainer;
int ando_trainer_create(int32_t device, ando_trainer** out);
int ando_trainer_destroy(ando_trainer* tr);
/* Adds the system A x = b (same BSR layout as ando_pcg); returns its index. */
int ando_trainer_add_system(ando_trainer* tr, int32_t num_rows, const int32_t* row_ptr, const int32_t* cols,
                            const float* values, const float* b, int32_t* index);
int ando_trainer_init_weights(ando_trainer* tr, uint64_t seed);
int ando_trainer_set_weights(ando_trainer* tr, const float* weights, int32_t count);
int ando_trainer_get_weights(ando_trainer* tr, float* weights, int32_t count);
/* Forward pass on one system: P (9 per row) and y = P b (the warm start is
 * alpha y with alpha = b^T y / y^T A y). */
int ando_trainer_forward(ando_trainer* tr, int32_t system, float* P, float* y);
/* Loss L = -(b^T y)^2 / ((y^T A y)(b^T D^-1 b)) (the A-norm error left by
 * the warm start, constant terms dropped) and its gradient (nullable). */
int ando_trainer_loss_grad(ando_trainer* tr, int32_t system, double* loss, float* grad);
/* steps Adam updates on systems drawn by a SplitMix64 stream from seed;
 * losses (nullable) receives the loss of each step's system before its update. */
int ando_trainer_train(ando_trainer* tr, int32_t steps, float lr, uint64_t seed, double* losses);
/* PCG on one system: learned = 0 plain block-Jacobi from x0 = 0, 1 learned
 * preconditioner and warm start, 2 block-Jacobi from the learned warm start
 * (no fallback). */
int ando_trainer_solve(ando_trainer* tr, int32_t system, int32_t learned, float tol, int32_t max_iters,
                       int32_t* iterations, int32_t* converged);

#ifdef __cplusplus
}
#endif

#endif
