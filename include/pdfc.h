/*
 * pdfc.h — C ABI of the MI355X projective-dynamics step with dry frictional
 * contact ("pdfc").
 *
 * One pdfc_sim owns one cloth scene on the current HIP device: the constant
 * system matrix K = M + dt^2 A^T A (factored once on the host, FP64), the
 * device-resident state (q, v), the per-step contact sets and all the
 * kernels' buffers. Every entry point returns PDFC_OK (0) or a negative
 * status; the message of the last failure on the calling thread is returned
 * by pdfc_last_error(). No C++ exception crosses this boundary.
 *
 * Reference interfaces replaced (file:line in guibquatre/ClimateNet):
 *   pdfc_create      getSimulation()            src/simulation_building.cpp:80-98
 *                    SimulationSpeed ctor       src/physics/SimulationSpeed.cpp:29-60
 *                    PhysicScene::finalize      src/physics/PhysicScene.cpp:88-91
 *   pdfc_step        SimulationBase::step()     src/physics/SimulationBase.cpp:103-173
 *   pdfc_get_state   PhysicScene::getGeneralizedPositions/Speeds
 *                                               src/physics/PhysicScene.cpp:228-305
 *   pdfc_set_state   PhysicScene::setGeneralizedPositions/Speeds
 *                                               src/physics/PhysicScene.cpp:268-314
 *   pdfc_get_contacts  m_collisions_infos + m_collisions_local_forces
 *                                               src/physics/SimulationFrictionEstimation.cpp:350-390
 *   pdfc_get_self_contacts  m_self_collisions_infos + m_self_collisions_local_forces
 *                                               src/physics/SimulationFrictionEstimation.cpp:437-540
 *   pdfc_get_iteration_stats  updateIterationData / Alart-Curnier norms
 *                                               src/physics/SimulationFrictionEstimation.cpp:590-612,720-753
 *   pdfc_get_phase_stats  exportStats timers     src/physics/SimulationFrictionEstimation.cpp:755-771
 *
 * Layout conventions at the boundary: positions and velocities are the
 * reference's generalized vectors (Eigen::VectorXd of size 3n, x,y,z of
 * vertex i at 3i..3i+2). Vertex indices are the scene's (OBJ) indices.
 */
#ifndef PDFC_H
#define PDFC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdfc_sim pdfc_sim;

#define PDFC_OK 0
#define PDFC_ERR_INVALID (-1)   /* bad argument or scene (std::invalid_argument in the reference) */
#define PDFC_ERR_DEVICE (-2)    /* HIP runtime failure */
#define PDFC_ERR_NUMERIC (-3)   /* factorisation failed (K not SPD) */
#define PDFC_ERR_CAPACITY (-4)  /* output buffer too small; *n holds the size needed */

enum pdfc_obstacle_type {
    PDFC_OBSTACLE_SPHERE = 0, /* StaticSphere: src/physics/StaticSphere.cpp:35-47 */
    PDFC_OBSTACLE_PLANE = 1,  /* StaticPlane:  src/physics/StaticPlane.cpp:57-73 */
    PDFC_OBSTACLE_MESH = 2,   /* StaticMesh (one frame, src/physics/StaticMesh.cpp:24-56) or MovingMesh
                                 (key frames, src/physics/MovingMesh.cpp:28-253); detection
                                 checkMeshCollision, src/geometry/geometry_util.cpp:32-92 */
    PDFC_OBSTACLE_BOX = 3,    /* StaticAxisAlignedBox: src/physics/StaticAxisAlignedBox.cpp:41-60 */
    PDFC_OBSTACLE_ROTATING_SPHERE = 4  /* RotatingSphere: src/physics/RotatingSphere.cpp:28-93 (a fixed
                                 sphere whose surface turns: contact speed omega x (contact - centre)) */
};

enum pdfc_precision {
    PDFC_FP64 = 0,           /* substitution and local step in FP64 */
    PDFC_FP32_REFINE64 = 1   /* FP32 substitution, FP64 residual + iterative refinement */
};

enum pdfc_warm_start {
    PDFC_WARM_PLAIN = 0,     /* the reference: r^0 is the cone projection of the contact force at v^0 */
    PDFC_WARM_PREVIOUS = 1,  /* r^0 = the previous step's reaction of the same (obstacle, vertex) contact,
                                plain where there was none */
    PDFC_WARM_LEARNED = 2    /* r^0 = the previous step's reaction plus the change the contact MLP predicts
                                (pdfc_set_mlp); the step restarts with the plain warm start when the
                                friction residual stagnates */
};

/* Contact MLP (learned warm start; no counterpart in the reference). */
#define PDFC_MLP_FEATURES 12
#define PDFC_MLP_OUTPUTS 3
#define PDFC_MLP_PARAMS 10448     /* FP32 master parameters, layout in climatenet_amd/csrc/mlp.hpp */
#define PDFC_MLP_NORM 27          /* feature mean[12], feature std[12], output scale[3] */

typedef struct pdfc_obstacle {
    int32_t type;            /* enum pdfc_obstacle_type */
    int32_t material;        /* "Material Identifier" */
    double point[3];         /* sphere centre, or a point of the plane ("Position" / "Point") */
    double normal[3];        /* plane normal ("Normal"), normalised by pdfc_create; unused for spheres */
    double radius;           /* sphere radius ("Radius") */
    /* triangle-mesh obstacles only (the data are copied by pdfc_create): */
    double tolerance;        /* "Threshold" (static) / "Collision tolerance" (moving) */
    size_t n_mesh_vertices;
    size_t n_mesh_triangles;
    const int32_t* mesh_triangles;   /* 3 per triangle */
    size_t n_frames;                 /* 1: static mesh; >= 2: key frames of a moving mesh */
    const double* mesh_frames;       /* n_frames x 3 n_mesh_vertices positions */
    const double* frame_times;       /* n_frames, increasing ("Time Between Frame" multiples); NULL if static */
    /* axis-aligned boxes only (point is "Position", the corner with the smallest coordinates): */
    double size[3];                  /* "Size", edge lengths along x, y, z */
    /* rotating spheres only (point and radius as for a sphere): */
    double rotation_axis[3];         /* "Rotation Axe" (any length but zero) */
    double angular_speed;            /* "Radial Speed", rad/s */
    double rotation_start;           /* "Rotation Start Time", s: the surface turns from the first step that
                                        begins once this much time has been stepped */
} pdfc_obstacle;

typedef struct pdfc_scene {
    size_t n_vertices;
    const double* rest_positions;    /* 3n, the "Obj filename" geometry (rest state) */
    const double* positions;         /* 3n or NULL: "Current state obj filename" geometry */
    size_t n_triangles;
    const int32_t* triangles;        /* 3T vertex indices, consistent winding */
    double area_density;             /* "Area Density" */
    double bending;                  /* "Bending" stiffness k_b (row weight sqrt(k_b)) */
    double stretching;               /* "Stretching" stiffness k_s (row weight sqrt(k_s)) */
    int32_t material;                /* mesh "Material Identifier" */
    size_t n_obstacles;
    const pdfc_obstacle* obstacles;
    size_t n_materials;
    const double* friction;          /* n_materials x n_materials, row-major ("Friction coefficients") */
    /* AttachmentConstraint rows after the mesh's constraints (PhysicScene::addConstraint,
       src/physics/PhysicScene.cpp:82-86; src/physics/AttachmentConstraint.cpp:23-55): row weight w
       on the vertex, projection w * position. No JSON key in the reference. */
    size_t n_attachments;
    const int32_t* attachment_vertices;  /* n_attachments vertex indices */
    const double* attachment_positions;  /* 3 n_attachments */
    const double* attachment_weights;    /* n_attachments, > 0 */
} pdfc_scene;

typedef struct pdfc_params {
    double time_step;                /* "Timestep" */
    int32_t local_global_iterations; /* "LocalGlobal" */
    int32_t self_collision;          /* "Self-collision" */
    double self_collision_tolerance; /* "Self-collision tolerance" */
    double gravity[3];               /* "Gravity" */
    double air_damping;              /* "Air damping" (optional, 0 when absent) */
    int32_t precision;               /* enum pdfc_precision */
    int32_t refinement_sweeps;       /* FP32_REFINE64 only: corrections after the first FP32 solve */
    int32_t warm_start;              /* enum pdfc_warm_start */
    int32_t record_iterations;       /* nonzero: keep the per-iteration friction residual (costs a sync-free reduction per iteration) */
    double residual_tolerance;       /* > 0: iterate until the mean friction (Alart-Curnier) residual of an
                                        iteration is below it, at most max_iterations; steps without contacts
                                        run local_global_iterations. 0: the reference's fixed count */
    int32_t max_iterations;          /* cap of the adaptive mode */
    int32_t use_graph;               /* nonzero: where a step needs no host round trip (no self-collision,
                                        fixed iteration count, plain warm start, no kernel timing) it is
                                        captured once as a HIP graph and replayed */
    int32_t warm_hold;               /* iterations that use the warm-start reaction instead of the cone
                                        projection (<= 1: iteration 0 only); in the adaptive mode the hold
                                        ends early once the friction residual stops decreasing */
    int32_t reserved;
} pdfc_params;

typedef struct pdfc_contact {
    int64_t vertex;                  /* CollisionInfo::vertex_index */
    int32_t obstacle;                /* index into pdfc_scene.obstacles */
    int32_t active;                  /* 0: take-off in the last iteration (r = 0) */
    double normal[3];                /* CollisionInfo::normal */
    double point[3];                 /* CollisionInfo::contact_point */
    double mu;                       /* CollisionInfo::friction_coefficient */
    double r_local[3];               /* reaction in the contact basis [n, t, b] (m_collisions_local_forces) */
    double speed[3];                 /* CollisionInfo::speed: obstacle velocity at the contact point */
} pdfc_contact;

typedef struct pdfc_self_contact {
    int64_t vertex;                  /* SelfCollisionInfo::vertex_index */
    int64_t partner;                 /* the face vertex paired with it (largest barycentric coordinate) */
    int64_t face[3];                 /* SelfCollisionInfo::face_indices */
    int32_t level;                   /* computation level (BFS order) */
    int32_t active;
    double normal[3];
    double mu;
    double r_local[3];               /* m_self_collisions_local_forces (mass-weighted) */
} pdfc_self_contact;

typedef struct pdfc_candidate {      /* one self-collision candidate surviving the exact test */
    int64_t vertex;
    int64_t triangle;
    double barycentric[3];
} pdfc_candidate;

typedef struct pdfc_info {
    int64_t n_vertices, n_triangles, n_springs, n_bending;
    int64_t nnz_K;                   /* A^T A / K non-zeros (full symmetric pattern) */
    int64_t nnz_L;                   /* non-zeros of the Cholesky factor (supernodal, dense blocks) */
    int64_t nnz_F;                   /* stored entries of the substitution operators (per copy) */
    int64_t n_supernodes;
    int64_t n_levels;                /* supernodal elimination-tree levels */
    int64_t n_contacts;              /* obstacle contacts of the last step */
    int64_t n_self_contacts;         /* self contacts of the last step */
    int64_t n_self_candidates;       /* exact-test survivors of the last detection (after pdfc_detect: all of
                                        them; after pdfc_step: those that pass the pairing pre-filter) */
    int64_t n_self_levels;
    int64_t max_front;               /* largest supernode (rows) */
    double factor_seconds;           /* host symbolic + numeric factorisation time */
    double bytes_per_iteration;      /* algorithmic HBM bytes of one local-global iteration */
    double bytes_solve;              /* of which the two substitutions */
    double bytes_local;              /* local-step kernel, per launch */
    double bytes_forward;            /* one forward sweep (all levels) */
    double bytes_backward;           /* one backward sweep (all levels) */
    int64_t n_below_rows;            /* total update-vector rows (sum of |R_s|) */
    double bytes_top;                /* dense top block of the substitution, per solve */
    int64_t top_m;                   /* rows of the dense top block (0: none) */
    int64_t n_levels_swept;          /* supernode levels swept per substitution (below the top block) */
    int32_t waves_forward[32];       /* waves per task chosen for each forward level (autotuned) */
    int32_t waves_backward[32];      /* same for the backward levels (root side first) */
    double autotune_ms;              /* time spent choosing them at creation */
    int32_t leaf_size;               /* largest leaf supernode of the dissection */
    int32_t top_chunks;              /* reduction split of the top-block matvec (autotuned); 0 when the
                                        symmetric tile products are used, or without a top block */
    double lhs_seconds;              /* host time to build topology, masses, constraints and A^T A (K) */
    int64_t steps_done;              /* steps taken since creation (restored by pdfc_load_checkpoint) */
    int32_t top_fold;                /* 1: the top block's ordered finish runs in the first backward level */
    int32_t launches_forward;        /* kernel launches per forward sweep (levels, or ranges of merged levels) */
    int32_t launches_backward;       /* same for the backward sweep */
    int32_t parts_forward[32];       /* levels per forward launch in sweep order (1: a level kernel; more: one
                                        launch, a workgroup per subtree of the range; autotuned) */
    int32_t parts_backward[32];      /* same for the backward sweep (root side first) */
    int32_t split_backward[32];      /* workgroups sharing each backward task's reduction, per backward level
                                        (1: one; 2 or 4: partial sums finished by the next level; autotuned) */
    int32_t top_pregather;           /* 1: the symmetric tiles read b_T gathered by a launch of its own */
} pdfc_info;

typedef struct pdfc_phase_stats {    /* device time of the last pdfc_step call, microseconds */
    double step_us;
    double detection_us;             /* host time per step of the self-collision round trip and pairing
                                        (including the wait for the device's detection and earlier work) */
    double iterations_us;            /* device time per step of its local-global iterations and epilogue */
    int64_t steps;
    int64_t iterations;              /* local-global iterations executed, all steps of the call */
    int64_t iterations_to_tolerance; /* adaptive mode: sum over steps with contacts of the first iteration
                                        count whose residual is below the tolerance (max_iterations when never) */
    int64_t contact_steps;           /* steps of the call with at least one contact */
    int64_t converged_steps;         /* of which reached the tolerance */
    int64_t fallbacks;               /* learned warm starts abandoned for the plain one */
    double collision_us;             /* device time per step from its start to its first iteration (prologue,
                                        obstacle and self detection, pairing); steps replayed from a HIP
                                        graph are not timed (then 0, and iterations_us 0) */
    int64_t timed_steps;             /* steps of the call that collision_us and iterations_us average over */
    double pairing_us;               /* host time per step of the self-contact pairing, level ordering and
                                        upload, after the detection results arrived (part of detection_us,
                                        which also waits for the device) */
    double order_us;                 /* of pairing_us: the level ordering (BFS) */
    double upload_us;                /* of pairing_us: packing the self contacts and enqueuing their upload */
    int64_t contacts_sum;            /* obstacle contacts summed over the call's steps (-1 when not known on
                                        the host: steps without self-collision, adaptive mode or warm data) */
    int64_t self_contacts_sum;       /* self contacts summed over the call's steps */
    int64_t self_levels_max;         /* most self-contact levels of one step in the call */
} pdfc_phase_stats;

typedef struct pdfc_kernel_times {  /* device time averaged over an instrumented run: kernel durations from
                                        the start and end events attached to each launch (hipExtLaunchKernel),
                                        sweeps also as event-pair intervals */
    double local_us;                 /* local step + RHS + obstacle friction kernel, per launch */
    double forward_us;               /* one forward-substitution level launch */
    double backward_us;              /* one backward-substitution level launch */
    double top_us;                   /* dense top block (gather, tiles or matvec, finish) per solve */
    double forward_sweep_us;         /* one forward sweep, one event pair around all its launches (gaps included) */
    double backward_sweep_us;
    int64_t local_launches, forward_launches, backward_launches, top_launches;  /* top: solves */
    double self_us;                  /* self-contact friction of one iteration (all its level launches) */
    int64_t self_iterations;         /* iterations with self contacts that self_us averages over */
} pdfc_kernel_times;

/* Build the simulation: topology, masses, constraints, K, host FP64 factorisation, device upload. */
int pdfc_create(const pdfc_scene* scene, const pdfc_params* params, pdfc_sim** out);

/* Batched instances (no counterpart in the reference, which runs one scene per process): count
 * independent scenes advanced together, each step one stream-ordered sequence of kernels over all
 * of them, K factored as one block-diagonal matrix (one elimination tree per scene). params apply
 * to every scene; self_collision must be 0. Vertices of scene k follow those of scenes 0..k-1 in
 * every per-vertex array of the other entry points, obstacles likewise; a vertex meets only the
 * obstacles of its own scene, with that scene's friction table. */
int pdfc_create_batch(const pdfc_scene* scenes, size_t count, const pdfc_params* params, pdfc_sim** out);

/* Advance nsteps time steps (SimulationBase::step() nsteps times). State stays on the device.
 * Returns PDFC_ERR_NUMERIC when a position or velocity became non-finite during the call (the
 * state is left as computed; the reference steps on silently). */
int pdfc_step(pdfc_sim* sim, int nsteps);

/* One step in the reference's phases, for a caller that drives SimulationBase::step()'s
 * protected hooks (include/physics/SimulationBase.hpp:123-131; SimulationBase.cpp:103-173):
 *   pdfc_step_begin    initializeStep + baseUpdateColisions + updateCollision: per-step constant,
 *                      obstacle and self-collision detection, pairing, levels (returns once the
 *                      pairing is done);
 *   pdfc_step_iterate  buildIteration + solve, n times: local step, friction, substitution
 *                      (enqueued, asynchronous);
 *   pdfc_step_end      updateScene: the new state; returns when the step is done.
 * pdfc_step_begin, pdfc_step_iterate(local_global_iterations), pdfc_step_end launches the kernels
 * of pdfc_step(sim, 1) in the same order (bit-identical states). Fixed iteration counts only
 * (residual_tolerance 0: PDFC_ERR_INVALID otherwise); any count >= 1 between begin and end. While
 * a step is open, pdfc_step, pdfc_set_state, pdfc_detect, pdfc_solve, pdfc_set_warm_reactions and
 * the checkpoint calls fail with PDFC_ERR_INVALID. */
int pdfc_step_begin(pdfc_sim* sim);
int pdfc_step_iterate(pdfc_sim* sim, int n);
int pdfc_step_end(pdfc_sim* sim);

/* Copy the generalized positions / velocities (3n each) to host memory; either may be NULL. */
int pdfc_get_state(pdfc_sim* sim, double* q, double* v);

/* Overwrite the state (3n each; either may be NULL). Used for identical-state parity and checkpoints.
 * PDFC_ERR_INVALID, with the state untouched, if a value is not finite. */
int pdfc_set_state(pdfc_sim* sim, const double* q, const double* v);

/* Binary checkpoint of the simulation state (no counterpart in the reference, whose
 * "Current state obj filename" restores positions only, as float, with zero velocity):
 * positions and velocities (FP64, OBJ vertex order), the obstacle clocks (moving meshes:
 * time, key frame and current positions; rotating spheres: time before the rotation), the
 * previous-step reactions kept for the PREVIOUS/LEARNED warm starts and the step count,
 * behind a format tag and version and followed by a 64-bit FNV-1a checksum. A simulation
 * built from the same scenes and parameters continues from it as the saved one would: bit
 * for bit with the same substitution schedule (the same simulation, or PDFC_TUNE pinned), to
 * rounding otherwise (the autotuner may choose another summation order).
 * save: with buf NULL or cap too small, returns PDFC_ERR_CAPACITY and *n = the size needed.
 * load: PDFC_ERR_INVALID, with the state untouched, if the buffer is not a checkpoint of a
 * simulation of this shape (tag, version, vertex/obstacle/mesh counts, size, checksum). */
int pdfc_save_checkpoint(pdfc_sim* sim, void* buf, size_t cap, size_t* n);
int pdfc_load_checkpoint(pdfc_sim* sim, const void* buf, size_t n);

/* Obstacle contacts of the last step, in the reference's order (obstacle, then vertex). */
int pdfc_get_contacts(pdfc_sim* sim, pdfc_contact* out, size_t cap, size_t* n);

/* Self contacts of the last step, in detection order (vertex ascending). */
int pdfc_get_self_contacts(pdfc_sim* sim, pdfc_self_contact* out, size_t cap, size_t* n);

/* Self-collision candidates of the last step's detection, sorted by (vertex, triangle). */
int pdfc_get_self_candidates(pdfc_sim* sim, pdfc_candidate* out, size_t cap, size_t* n);

/* Run only the detection of the current state (no time step) so candidates/contacts can be read. */
int pdfc_detect(pdfc_sim* sim);

/* Per-iteration mean and max Alart-Curnier norm over all contacts of the last step
 * (requires record_iterations). ac_mean/ac_max may be NULL. */
int pdfc_get_iteration_stats(pdfc_sim* sim, double* ac_mean, double* ac_max, size_t cap, size_t* n_iter);

int pdfc_get_phase_stats(pdfc_sim* sim, pdfc_phase_stats* out);

/* Instrumentation (no counterpart in the reference; measurement only): when enabled, every
 * launch of the local kernel and of each substitution level is bracketed by HIP events on the
 * simulation's stream, and pdfc_step synchronises after each step to accumulate the times.
 * Enabling resets the accumulated times. */
int pdfc_set_kernel_timing(pdfc_sim* sim, int enable);
int pdfc_get_kernel_times(pdfc_sim* sim, pdfc_kernel_times* out);
int pdfc_get_info(pdfc_sim* sim, pdfc_info* out);

/* Solve K x = b for 3 right-hand sides (3n, generalized layout) with the device substitution
 * in the simulation's precision mode. Exposed for parity tests of the global step. */
int pdfc_solve(pdfc_sim* sim, const double* b, double* x);

void pdfc_destroy(pdfc_sim* sim);
const char* pdfc_last_error(void);

/* ---- learned warm start. No counterpart in the reference: a per-contact impulse guess for
 * iteration 0 of the local-global loop (the reference computes r^0 as the cone projection of the
 * contact force at v^0, SimulationFrictionEstimation.cpp:350-423).
 *
 * Features of an obstacle contact (vertex i, basis B = [n t b], mass m_i), evaluated at the start
 * of the step: [0..2] relative velocity B^T(v^0_i - u_obstacle); [3..5] contact normal (world);
 * [6..8] previous-step reaction of the same (obstacle, vertex) pair per unit mass, in B; [9] mu;
 * [10] gap (signed distance of q_n,i to the obstacle surface); [11] normal component of the
 * contact force at v^0 per unit mass (the plain warm start's own estimate). The network predicts
 * the change of the converged reaction per unit mass in B from the previous step's (features
 * 6..8); the step uses m_i times (previous + predicted change), projected on the cone. */
typedef struct pdfc_mlp pdfc_mlp;

/* 12 -> 64 -> 64 -> 64 -> 3, SiLU, Glorot-uniform weights from seed, identity normalisation. */
int pdfc_mlp_create(pdfc_mlp** out, uint64_t seed);
void pdfc_mlp_destroy(pdfc_mlp* mlp);
/* params: PDFC_MLP_PARAMS floats (NULL keeps them); norm: PDFC_MLP_NORM floats (NULL keeps them).
 * Resets the optimiser state. */
int pdfc_mlp_set_params(pdfc_mlp* mlp, const float* params, const float* norm);
int pdfc_mlp_get_params(pdfc_mlp* mlp, float* params, float* norm);
/* y[rows x 3] = scale * MLP((x - mean) / std) on the device (BF16 MFMA, FP32 accumulation). */
int pdfc_mlp_forward(pdfc_mlp* mlp, const float* x, size_t rows, float* y);
/* Gradient of L = mean over rows and outputs of (MLP(x) - t / scale)^2 w.r.t. the parameters. */
int pdfc_mlp_gradient(pdfc_mlp* mlp, const float* x, const float* t, size_t rows, float* grad, float* loss);
/* Allocation-free gradient. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training or serving loops, harnesses
 * that forbid hipMalloc). The same gradient and mean loss as pdfc_mlp_gradient, by the same
 * kernels in the same summation order, with every operand in DEVICE memory: x [rows x 12] and
 * t [rows x 3] in, grad [PDFC_MLP_PARAMS] and loss [1] out. The per-tile partial gradients live in
 * a fixed array of the library, one per GPU (PDFC_MLP_NOALLOC_MAX_ROWS / 16 tiles, 5.35 MB), held
 * until the process ends; more rows are rejected with PDFC_ERR_INVALID. Calls no hipMalloc,
 * hipFree or other allocating routine. Runs on the network's stream and returns when it is done.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (several threads or
 * networks at once) race. */
#define PDFC_MLP_NOALLOC_MAX_ROWS 2048
int pdfc_mlp_gradient_noalloc(pdfc_mlp* mlp, const float* x, const float* t, size_t rows, float* grad, float* loss);
/* Adam (beta 0.9 / 0.999) on shuffled mini-batches of batch rows (multiple of 16) for epochs epochs;
 * loss_per_epoch (optional, epochs floats) gets each epoch's mean mini-batch loss. */
int pdfc_mlp_train(pdfc_mlp* mlp, const float* x, const float* t, size_t rows, int epochs, int batch, float lr,
                   uint64_t seed, float* loss_per_epoch);
/* Warm start of the next step only, from given reactions: a contact of that step whose (vertex,
 * obstacle) matches one of contacts[0..n) starts from its r_local (expressed in the basis of its
 * normal), projected on the cone, for the first hold iterations; other contacts start plain.
 * Used to measure what an exact warm start (the converged reactions of the same step) gains. */
int pdfc_set_warm_reactions(pdfc_sim* sim, const pdfc_contact* contacts, size_t n, int hold);
/* Attach a network to a simulation (not owned; must outlive its use; NULL detaches). */
int pdfc_set_mlp(pdfc_sim* sim, pdfc_mlp* mlp);
/* Obstacle contacts of the last step (pdfc_get_contacts order): features[n x 12] and the final
 * reaction per unit mass in the contact basis, targets[n x 3]. Available when warm_start is not
 * PLAIN or residual_tolerance > 0. */
int pdfc_get_warm_data(pdfc_sim* sim, float* features, float* targets, size_t cap, size_t* n);

#ifdef __cplusplus
}
#endif
#endif /* PDFC_H */
