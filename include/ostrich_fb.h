/*
 * ostrich_fb.h -- C ABI of the MI355X (gfx950) Fisher-Burmeister Newton contact solver.
 *
 * This is the drop-in boundary for Ostrich's solver plugin
 * (`OstrichEngine(SolverBase)`, src/ostrich/core/engine.py:14, constructed by
 * `OstrichEngineConfig.create_engine`, src/ostrich/core/engine_config.py:527-557).
 * The Python class `ostrich_amd.OstrichEngine` binds these entry points with ctypes;
 * every pointer named "device" is a HIP device pointer owned by the caller and valid
 * for the duration of the call.  Scratch memory is owned by the solver handle and is
 * allocated in ofb_create only, so ofb_step allocates nothing and can be captured in a
 * hipGraph.
 *
 * Conventions (identical to Newton's, as read by the reference):
 *   transform   = px py pz qx qy qz qw            (7 x f32, quaternion xyzw)
 *   spatial     = linear(3) angular(3)            (6 x f32, world frame;
 *                 src/ostrich/constraints/dynamics_constraint.py:40)
 *   body arrays are world-major: index = world * bodies_per_world + body
 *   joint types: core/model.py:88-99; dof modes: core/types.py:4-18
 *
 * Return codes: 0 = ok, negative = invalid argument / shape / launch failure; the
 * message is available from ofb_last_error().  Non-convergence is not an error: it is
 * reported by ofb_get_stats (src/ostrich/core/base_engine.py:42-80 has no error path).
 */
#ifndef OSTRICH_FB_H
#define OSTRICH_FB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OFB_JOINT_PRISMATIC = 0,
  OFB_JOINT_REVOLUTE = 1,
  OFB_JOINT_BALL = 2,
  OFB_JOINT_FIXED = 3,
  OFB_JOINT_FREE = 4,
  OFB_JOINT_DISTANCE = 5,
  OFB_JOINT_D6 = 6,
  OFB_JOINT_CABLE = 7
};

enum { OFB_DOF_NONE = 0, OFB_DOF_TARGET_POSITION = 1, OFB_DOF_TARGET_VELOCITY = 2 };

/* linear.preconditioner_type, core/engine_config.py:137 */
enum { OFB_PRECOND_JACOBI = 0, OFB_PRECOND_PER_BODY_PAIR = 1 };

/* relaxation.friction / relaxation.contact, core/engine_config.py:243-244 */
enum { OFB_FRICTION_CONE = 0, OFB_FRICTION_BILATERAL = 1, OFB_FRICTION_BILATERAL_PATCH = 2 };
enum { OFB_CONTACT_COMPLEMENTARITY = 0, OFB_CONTACT_EQUALITY = 1 };

/* contacts.reduction.policy (collision/config.py:13-56): per (body0, body1) pair, keep
 * the K deepest contacts (TOP_K, collision/top_k.py), a farthest-point sample seeded by
 * the deepest (FPS, collision/fps.py), or the K deepest cluster representatives
 * (CLUSTER, collision/cluster.py).  Applied to the packed contacts before the warm start,
 * with the step-start poses; survivors keep their packed order. */
enum { OFB_REDUCE_NONE = 0, OFB_REDUCE_TOP_K = 1, OFB_REDUCE_FPS = 2, OFB_REDUCE_CLUSTER = 3 };

/* warm_start.method (collision/warm_start.py): match each contact to the previous step's
 * nearest contact of its body pair (position_match, :84-519), or spread each body pair's
 * previous total normal impulse evenly over the pair's current contacts (pair_aggregate,
 * :896-1010). */
enum { OFB_WS_POSITION_MATCH = 0, OFB_WS_PAIR_AGGREGATE = 1 };

/* Termination of the Newton and PCR loops.
 * GLOBAL:    the reference's batch-global stop flag (base_engine.py:42-80,
 *            pcr_solver.py:84-117): every world iterates until all worlds converge.
 * PER_WORLD: each world stops on its own criterion and is retired (fused path). */
enum { OFB_TERM_GLOBAL = 0, OFB_TERM_PER_WORLD = 1 };

/* Execution path. FUSED: one persistent launch, one wavefront per world, rows held in
 * registers.  PLAIN: one kernel per reference kernel family, data in HBM. */
enum { OFB_PATH_FUSED = 0, OFB_PATH_PLAIN = 1 };

/* Model description; mirrors OstrichModel (src/ostrich/core/model.py:171-342).
 * All worlds must share one topology (model.py:200-203). */
typedef struct ofb_model {
  int32_t num_worlds;
  int32_t bodies_per_world;
  int32_t joints_per_world;
  int32_t dofs_per_world;
  int32_t shapes_per_world;   /* S: shapes owned by each world */
  int32_t num_global_shapes;  /* G: shapes shared by every world (ground, heightfield) */
  float gravity[3];
  /* bodies [W*B] */
  const float* body_mass;
  const float* body_inv_mass;
  const float* body_inertia;     /* [W*B*9] row-major, body frame */
  const float* body_inv_inertia; /* [W*B*9] */
  const float* body_com;         /* [W*B*3] body frame */
  /* joints [W*J] */
  const int32_t* joint_type;
  const int32_t* joint_parent; /* world-local body index, -1 = world */
  const int32_t* joint_child;
  const int32_t* joint_qd_start; /* world-local dof index */
  const int32_t* joint_enabled;
  const float* joint_X_p;        /* [W*J*7] */
  const float* joint_X_c;        /* [W*J*7] */
  const float* joint_compliance; /* [W*J], < 0 selects config.compliance_joint */
  /* dofs [W*D] */
  const float* joint_axis; /* [W*D*3] */
  const int32_t* joint_dof_mode;
  const float* joint_target_ke;
  const float* joint_target_kd;
  /* shapes [W*(S+G)]: per-world columns [0,S) then the globals [S,S+G) */
  const int32_t* shape_body; /* world-local body, -1 = static */
  const float* shape_mu;
  const float* shape_friction_axis; /* [W*(S+G)*3] */
  const float* shape_mu_perp;
  const float* shape_mu_stiction_scale;
  const float* shape_v_stribeck;
  const float* shape_stribeck_lateral_only;
  /* Newton flat shape index -> world, [G + W*S]; globals hold -1 */
  const int32_t* shape_world;
} ofb_model;

/* Solver configuration; mirrors OstrichEngineConfig (core/engine_config.py:103-577). */
typedef struct ofb_config {
  int32_t nr_max_iters;          /* nr.max_iters */
  int32_t nr_backtrack_min_iter; /* nr.backtrack_min_iter */
  float nr_atol;                 /* nr.atol on ||h||_2 */
  int32_t lin_max_iters;         /* linear.max_iters */
  float lin_tol;                 /* linear.tol */
  float lin_atol;                /* linear.atol */
  int32_t preconditioner;        /* OFB_PRECOND_* */
  float regularization;          /* linear.regularization */
  float compliance_joint;
  float compliance_contact;
  float compliance_friction;
  float contact_fb_eps_sq;       /* compliance.contact_fb_smooth_eps_sq */
  float contact_fb_alpha;        /* CONTACT_FB_ALPHA, contact_constraint.py:9 */
  float gyro_scale;              /* relaxation.gyro -> 1 or 0 */
  float friction_load_gate;      /* relaxation.friction_load_gate */
  int32_t friction_mode;         /* OFB_FRICTION_* */
  int32_t contact_mode;          /* OFB_CONTACT_* */
  int32_t max_contacts_per_world;/* contacts.max_per_world */
  int32_t termination;           /* OFB_TERM_* */
  int32_t seed_iterate;          /* warm_start.seed_iterate: start NR from the lambda seed */
  int32_t rigid_contact_max;     /* capacity of the flat contact arrays (Newton
                                    Contacts.rigid_contact_max); sizes the packing scratch */
  float adjoint_regularization;  /* adjoint.regularization: gamma added to C of active rows */
  int32_t adjoint_pose_vjp;      /* adjoint.pose_vjp: pull w^T dR/dq+ and w^T dR/dq- into the
                                    pose gradients */
  float adjoint_pose_eps;        /* central-difference step of that pull-back, in pose
                                    coordinates (m and quaternion components) */
  int32_t warm_start;            /* warm_start.enabled: seed each step from the previous
                                    step's contacts (acts only with seed_iterate, as in
                                    the reference) */
  int32_t warm_start_cold_gravity; /* warm_start.cold_gravity */
  int32_t warm_start_cold_impact;  /* warm_start.cold_impact */
  float warm_start_cold_friction_v;/* warm_start.cold_friction_v_threshold (<= 0: off) */
  int32_t linesearch;            /* linesearch.enabled (fused path only) */
  int32_t linesearch_conservative_count;   /* linesearch.conservative_step_count */
  int32_t linesearch_optimistic_count;     /* linesearch.optimistic_step_count */
  double linesearch_min_step;              /* linesearch.min_step (double: the candidate set
                                              is built in double, as the reference's numpy) */
  double linesearch_conservative_upper;    /* linesearch.conservative_upper_bound */
  double linesearch_optimistic_window;     /* linesearch.optimistic_window */
  int32_t pair_small_worlds;     /* fused path: worlds of at most 32 rows run two per
                                    wavefront on the assembled (dense) Schur operator;
                                    0: one world per wavefront, matrix-free operator */
  int32_t contact_reduction;     /* contacts.reduction.policy: OFB_REDUCE_* (collision/config.py:13) */
  int32_t contact_max_per_pair;  /* contacts.reduction.max_per_pair: K per (body0, body1) pair */
  float cluster_normal_dot;      /* contacts.reduction.cluster_normal_dot_thresh */
  float cluster_pos;             /* contacts.reduction.cluster_pos_thresh [m] */
  int32_t warm_start_method;     /* warm_start.method: OFB_WS_* */
  int32_t adjoint_friction_frozen; /* adjoint.friction_linearization == "frozen": the adjoint
                                      replaces each friction row's compliance by the stick /
                                      slip surrogate of adjoint/adjoint_friction.py:20-158
                                      (rows of contacts without normal impulse dropped) and
                                      leaves the friction rows out of w^T dR/dxi
                                      (base_engine.py:536-590, 704-711) */
  int32_t adjoint_soft_blending;   /* adjoint.soft_blending: sigmoid between the two
                                      compliances instead of the 0.9 threshold */
  float adjoint_soft_temperature;  /* adjoint.soft_blending_temperature */
  int32_t large_world_slots;       /* fused path, Jacobi: layout of the large worlds; 2 (or
                                      0): worlds of more than 128 rows on 2, 4 or 8
                                      wavefronts with two row slots per lane, the shortest
                                      solve of one world; 4 (the wide layout): 65-512 rows
                                      on two wavefronts with two to four slots per lane,
                                      513-1024 on four with four: about twice as many large
                                      worlds resident at once.  No reference counterpart (a
                                      launch layout); like pair_small_worlds a
                                      rounding-level difference */
} ofb_config;

typedef struct ofb_state_in {
  const float* body_q;  /* [W*B*7] */
  const float* body_qd; /* [W*B*6] */
  const float* body_f;  /* [W*B*6] */
} ofb_state_in;

typedef struct ofb_state_out {
  float* body_q;  /* [W*B*7] */
  float* body_qd; /* [W*B*6] */
} ofb_state_out;

typedef struct ofb_control {
  const float* joint_target_pos; /* [W*D] */
  const float* joint_target_vel; /* [W*D] */
} ofb_control;

/* Newton rigid contacts (flat, unbatched) as read by core/contacts.py:25-99.
 * normal points from shape0 to shape1 (Newton A-to-B); point0/point1 are in the frame
 * of the owning body (world frame for static shapes); margins become thicknesses. */
typedef struct ofb_contacts {
  const int32_t* count; /* [1] device */
  int32_t max_contacts;
  const float* point0;  /* [max*3] */
  const float* point1;  /* [max*3] */
  const float* normal;  /* [max*3] */
  const int32_t* shape0;
  const int32_t* shape1;
  const float* margin0;
  const float* margin1;
} ofb_contacts;

/* Opaque solver handle (engine state + scratch), one per OstrichEngine instance. */
typedef struct ofb_solver ofb_solver;
typedef struct ofb_mlp ofb_mlp;

/* Replaces OstrichEngineBase.__init__ (core/base_engine.py:112-237). Copies the model
 * into solver-owned device memory and allocates every scratch buffer. `path` selects
 * OFB_PATH_FUSED or OFB_PATH_PLAIN. Model pointers are device pointers. */
int ofb_create(const ofb_model* model, const ofb_config* config, int32_t path,
               ofb_solver** out_solver);

/* Releases everything ofb_create allocated. */
int ofb_destroy(ofb_solver* solver);

/* Number of constraint rows per world, N_c = N_j + N_ctrl + 3*C_max
 * (core/engine_dims.py:56-61), and the block offsets of the row layout
 * [joints | control | normals | friction] (engine_dims.py:68-86). */
int ofb_dims(const ofb_solver* solver, int32_t* n_c, int32_t* offset_ctrl, int32_t* offset_n,
             int32_t* offset_f);

/* Replaces OstrichEngine.step (core/engine.py:25-71): load_data (base_engine.py:289-335),
 * the Newton loop (_solve, base_engine.py:416-526), best-iterate backtracking
 * (core/backtracking.py:12-95) and the copy-out of body_q/body_qd. All pointers are
 * device pointers; `stream` is a hipStream_t (NULL = default stream). No host
 * synchronisation happens inside. */
int ofb_step(ofb_solver* solver, const ofb_state_in* state_in, const ofb_state_out* state_out,
             const ofb_control* control, const ofb_contacts* contacts, float dt, void* stream);

/* Per-world statistics of the last step, copied into caller device buffers (any may be
 * NULL): Newton iterations, summed PCR iterations, the picked merit ||h||^2, the picked
 * candidate index and a flag word (bit0 NR cap hit, bit1 contact overflow, bit2 block
 * factor failure, bit3 pair member overflow). Replaces the HDF5 history
 * (core/base_engine.py:239-247). */
int ofb_get_stats(const ofb_solver* solver, int32_t* nr_iters, int32_t* pcr_iters,
                  float* merit, int32_t* best_iter, int32_t* flags, void* stream);

/* Copies the picked iterate's constraint impulses lambda [W*N_c] (data._constr_force
 * after perform_backtracking) into a caller device buffer. */
int ofb_get_lambda(const ofb_solver* solver, float* lambda_out, void* stream);

/* Copies lambda_prev_iter [W*N_c] as the last step left it: the snapshot taken at the
 * start of its last Newton iteration (core/base_engine.py:441), which backtracking does
 * not restore.  With lambda it is what TrajectoryBuffer.save_step keeps per step
 * (simulation/trajectory_buffer.py:222-223) for the adjoint.  Fused path only; -2 when
 * the solver does not keep it (ofb_set_keep_lambda_prev). */
int ofb_get_lambda_prev(const ofb_solver* solver, float* lambda_prev_out, void* stream);

/* Whether the following steps write lambda_prev_iter for ofb_get_lambda_prev (default 1).
 * The reference keeps it in its engine data and allocates the adjoint's arrays only for
 * differentiable_simulation (core/base_engine.py:118-146); a caller that never runs the
 * adjoint sets 0 and the fused kernel skips that W*N_c*4-byte write per step.  Plays no
 * part in the forward solve. */
int ofb_set_keep_lambda_prev(ofb_solver* solver, int32_t keep);

/* Replaces the per-shape friction coefficients shape_mu [W*(S+G)] (OstrichModel
 * shape_material_mu, core/model.py) copied at ofb_create, e.g. between steps of a
 * gradient-based fit of mu.  Device pointer. */
int ofb_set_shape_mu(ofb_solver* solver, const float* shape_mu, void* stream);

/* Gradients flowing into one step's outputs: dL/dq+ [W*B*7] and dL/du+ [W*B*6]
 * (data.body_pose_grad / body_vel_grad loaded by TrajectoryBuffer.load_step,
 * simulation/trajectory_buffer.py:241-242). */
typedef struct ofb_adjoint_in {
  const float* body_q_grad;
  const float* body_qd_grad;
} ofb_adjoint_in;

/* Gradients with respect to the step's inputs.  Every pointer may be NULL; non-NULL
 * outputs are overwritten (not accumulated). */
typedef struct ofb_adjoint_out {
  float* body_q_grad;           /* dL/dq-       [W*B*7]  body_pose_prev.grad */
  float* body_qd_grad;          /* dL/du-       [W*B*6]  body_vel_prev.grad */
  float* body_f_grad;           /* dL/df        [W*B*6]  ext_force.grad */
  float* joint_target_pos_grad; /* [W*D] */
  float* joint_target_vel_grad; /* [W*D] */
  float* joint_target_ke_grad;  /* [W*D] */
  float* joint_target_kd_grad;  /* [W*D] */
  float* shape_mu_grad;         /* [W*(S+G)] per world copy of shape_mu */
  float* adjoint_lambda;        /* w_lambda [W*N_c], reference row layout (data.w.c) */
  float* adjoint_body;          /* w_u [W*B*6] (data.w.d_spatial) */
} ofb_adjoint_out;

/* Adjoint of one step (OstrichEngine.step_backward, core/base_engine.py:528-743) at the
 * state the forward step left: state_in = q-, u-, f of the step start; state_out = q+, u+
 * after it (read only); lambda / lambda_prev from ofb_get_lambda / ofb_get_lambda_prev of
 * that step; contacts = the step's contact set (re-packed here).  Rebuilds J, C, h at
 * that state, forms w_u0 = M^-1 f(dL/dq+, dL/du+), solves (J M^-1 J^T + C + reg) w_lambda
 * = J w_u0 with the configured PCR and preconditioner, sets w_u = w_u0 - M^-1 J^T w_lambda
 * and returns w^T dR/dxi for every input xi plus the explicit pose carry
 * (residual.py:176-406).  Per-world adjoint PCR counts and flags (bit 4: a friction row's
 * mu derivative has no closed form, that part of shape_mu_grad is missing) are read with
 * ofb_get_stats.  Fused path only. */
int ofb_step_backward(ofb_solver* solver, const ofb_state_in* state_in, const ofb_state_out* state_out,
                      const ofb_control* control, const ofb_contacts* contacts, const float* lambda,
                      const float* lambda_prev, const ofb_adjoint_in* grad_in, const ofb_adjoint_out* grad_out,
                      float dt, void* stream);

/* Forgets the warm starter's snapshot (ContactWarmStarter state, collision/warm_start.py:
 * 620-640), e.g. at an episode reset: the next step cold-starts every contact. */
int ofb_warm_start_reset(ofb_solver* solver, void* stream);

/* The warm starter's snapshot as one opaque device block (no reference counterpart: the
 * reference keeps it in engine buffers that a re-simulation overwrites).  A checkpointed
 * rollout saves it with each checkpoint and restores it before re-simulating a window, so
 * the re-simulated steps start Newton from the seeds the forward pass used.  _size gives
 * the block's bytes (0 when the warm start is off); _get / _set copy it device to device
 * on `stream` (caller-owned buffer of that size). */
int ofb_warm_start_state_size(const ofb_solver* solver, int64_t* bytes);
int ofb_warm_start_state_get(const ofb_solver* solver, void* dst, void* stream);
int ofb_warm_start_state_set(ofb_solver* solver, const void* src, void* stream);

/* Copies the Newton seed [W*N_c] the last step started from (warm starter output, or the
 * seed set with ofb_set_lambda_seed) - _constr_force after engine.py:55-59. */
int ofb_get_lambda_seed(const ofb_solver* solver, float* seed_out, void* stream);

/* ---- per-contact warm-start network (no counterpart in the reference) ----------------
 * 12 -> 64 -> 64 -> 3 ReLU MLP in BF16 with FP32 accumulation on MFMA (fb_mlp.hip).  x is
 * [n*12] features, y and target [n*3] (device pointers); parameters are [num_params] FP32
 * master weights (W1 16x64, b1, W2 64x64, b2, W3 64x16, b3; padding rows/columns zero). */
int ofb_mlp_create(ofb_mlp** out);
int ofb_mlp_destroy(ofb_mlp* net);
int ofb_mlp_num_params(void);
int ofb_mlp_set_params(ofb_mlp* net, const float* params, void* stream);
int ofb_mlp_get_params(const ofb_mlp* net, float* params, void* stream);
int ofb_mlp_forward(const ofb_mlp* net, const float* x, int n, float* y, void* stream);
/* gradient of mean_rows sum_c (y_c - target_c)^2 w.r.t. the parameters; loss may be NULL */
int ofb_mlp_gradient(ofb_mlp* net, const float* x, const float* target, int n, float* grad, float* loss,
                     void* stream);
/* ofb_mlp_gradient that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * training loops, harnesses that forbid hipMalloc).  ofb_mlp_gradient grows its per-block
 * gradient slots with hipMalloc when n grows; this one keeps them in a file-scope device
 * array of the library (at most 8192 rows: 3.2 MB that each GPU holds until the process
 * ends) and calls no hipMalloc / hipFree.  Same kernels, so the same bits as
 * ofb_mlp_gradient.  grad (NULL: the network's own buffer) and loss (may be NULL) are
 * device pointers.  Returns -2 for n > 8192.  Calls that overlap on one GPU share the
 * array and race: serialise them (one stream, or events). */
int ofb_mlp_gradient_noalloc(ofb_mlp* net, const float* x, const float* target, int n, float* grad, float* loss,
                             void* stream);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss */
int ofb_mlp_train_step(ofb_mlp* net, const float* x, const float* target, int n, float lr, float* loss,
                       void* stream);

/* Copies the merit line search's candidate step sizes (core/engine_data.py:15-39, the
 * set the solver evaluates when linesearch is enabled) into a host buffer of at least
 * conservative_count + optimistic_count floats; *count receives their number. */
int ofb_get_linesearch_steps(const ofb_solver* solver, float* steps, int32_t* count);

/* Attaches (net != NULL) or detaches the learned warm start (no counterpart in the
 * reference): with warm_start and seed_iterate on, each step the network predicts the
 * contact impulses from the warm starter's features and every world starts Newton from
 * whichever of the two seeds (network, previous step) has the lower initial residual
 * (stats flag bit 5: the network's).  The network is not owned.  Fused path only. */
int ofb_set_warm_start_net(ofb_solver* solver, const ofb_mlp* net);

/* Copies the learned warm start's per-contact features [W*C_max*12] of the last step and
 * their impulse scale m_eff |g| [W*C_max] (dataset logging).  Needs warm_start. */
int ofb_get_contact_features(const ofb_solver* solver, float* features, float* scale, void* stream);

/* Copies the packed per-world contact set of the last step: count [W] and, for each of
 * the W*C_max slots, shape0/shape1 (world-local columns), point0/point1/normal (after the
 * normal flip, core/contacts.py:80-83) and thickness0/thickness1. Any pointer may be NULL. */
int ofb_get_packed_contacts(const ofb_solver* solver, int32_t* count, int32_t* shape0,
                            int32_t* shape1, float* point0, float* point1, float* normal,
                            float* thickness0, float* thickness1, void* stream);

/* Sets the lambda [W*N_c] that the next ofb_step starts Newton from when
 * config.seed_iterate is on (core/engine.py:55-59); also the injection point of the
 * learned warm start. */
int ofb_set_lambda_seed(ofb_solver* solver, const float* lambda_seed, void* stream);

/* Duration in milliseconds of the last step's solve (the fused kernel, or the plain
 * path's Newton loop), measured with HIP events recorded on the step's stream around
 * those launches; contact packing is excluded. Blocks until that work has finished. */
int ofb_last_solve_ms(const ofb_solver* solver, float* ms);

/* Duration in milliseconds of the last step's paired-class fused kernel (worlds of at
 * most 32 rows, two per wavefront: the launch that holds most worlds of a typical batch),
 * from HIP events recorded on the step's stream directly around that launch.  No
 * reference counterpart (measurement hook).  Returns -2 when the last step launched no
 * such kernel (plain path, pairing off) or was captured into a graph, in which case no
 * step timing is recorded (ofb_last_solve_ms also returns -2). */
int ofb_last_kernel_ms(const ofb_solver* solver, float* ms);

/* Per-class timing of the fused path (no reference counterpart; measurement hook).  With
 * timing on, every eager step records HIP events around each row-class launch on the
 * stream it runs on (the step's stream or one of the solver's side streams, which run
 * beside it).
 * ofb_get_class_times fills, for each of the `count` classes (capacity >= 10), its row
 * capacity, the number of worlds the last step put in it and its duration in ms (-1 when
 * the class was not launched or the step was not timed); it blocks until the step has
 * finished.  ofb_get_class_shapes fills, per class, the row slots per lane and the lanes
 * per world of the kernel instance that class runs (lanes 32: two worlds per wavefront;
 * 64: one world per wavefront; 128, 256, 512: one world over 2, 4, 8 wavefronts of its
 * workgroup) and the largest row count of a world in that class (0: the class is not
 * launched); a world of n rows is in the first class whose largest row count is >= n.
 * -2 on the plain path. */
int ofb_set_class_timing(ofb_solver* solver, int32_t enabled);
int ofb_get_class_times(const ofb_solver* solver, int32_t capacity, int32_t* rows_cap, int32_t* worlds, float* ms,
                        int32_t* count);
int ofb_get_class_shapes(const ofb_solver* solver, int32_t capacity, int32_t* slots_per_lane,
                         int32_t* lanes_per_world, int32_t* max_rows, int32_t* count);

/* Launch settings of the fused path: the largest number of wavefronts per workgroup of a
 * row-class launch (1..16; the kernels' 256-thread bound caps it at 4, and the LDS limit
 * may lower it per class) and whether worlds of at
 * most 32 rows run two per wavefront.  No reference counterpart (Warp chooses its own
 * launch shapes); the autotuner (OstrichEngine.autotune) sets them from measurements.
 * The workgroup size leaves results bit-identical; pairing changes the width of the
 * lane-group reductions, a rounding-level difference.  Fused path only (-2 otherwise). */
int ofb_set_fused_launch(ofb_solver* solver, int32_t waves_per_block, int32_t pair_small_worlds);
int ofb_get_fused_launch(const ofb_solver* solver, int32_t* waves_per_block, int32_t* pair_small_worlds);

/* Layout of the large worlds of the fused Jacobi path: ofb_config.large_world_slots (2 or
 * 4; -2 otherwise), applied to the next step.  No reference counterpart; the autotuner
 * (OstrichEngine.autotune) chooses it with the launch settings.  Fused path only. */
int ofb_set_fused_layout(ofb_solver* solver, int32_t large_world_slots);
int ofb_get_fused_layout(const ofb_solver* solver, int32_t* large_world_slots);

/* One Adam step on a control trajectory params[steps][worlds][dofs] (device, f32), moments
 * m and v of the same shape (zero before the first step), for the dof slice
 * [dof_offset, dof_offset + num_dofs) of every (step, world): optim/trajectory_adam.py:
 * 6-53 and TrajectoryAdam.step (:105-131), with t the 0-based step counter.  Returns -2
 * on a slice outside the array. */
int ofb_trajectory_adam_step(float* params, const float* grad, float* m, float* v, int32_t steps,
                             int32_t worlds, int32_t dofs, int32_t dof_offset, int32_t num_dofs, float lr,
                             float beta1, float beta2, float eps, float clip_grad, int32_t t, void* stream);

/* Attaches (merit != NULL) or detaches caller device buffers [capacity][W] that each
 * fused step fills with its per-Newton-iteration history: ||h||^2 after iteration k
 * (the reference's candidates_res_norm_sq, core/history_group.py buffers logged by
 * logging/sim_logger.py) and the PCR iterations of iteration k (pcr_history_iter_count);
 * the first 32 iterations are recorded, later ones and iterations a world did not run
 * hold NaN and -1.  Fused path only. */
int ofb_set_iteration_history(ofb_solver* solver, float* merit, int32_t* pcr_iters, int32_t capacity);

/* ---- device-resident step logs (core/history_group.py:12-80, logging/sim_logger.py,
 * logging/dataset_logger.py) ----------------------------------------------------------
 * A log copies num_arrays device arrays of words_per_step[a] 4-byte elements into slot
 * *step_index of their [capacity, words] destination buffers (all device pointers, kept
 * by the caller); a step index outside [0, capacity) copies nothing.  The index is read
 * on the device and, with advance, incremented after the copy, so a capture records into
 * a graph. */
typedef struct ofb_log ofb_log;
int ofb_log_create(int32_t num_arrays, const void* const* src, void* const* dst, const int64_t* words_per_step,
                   int32_t capacity, ofb_log** out);
int ofb_log_capture(ofb_log* log, int32_t* step_index, int32_t advance, void* stream);
int ofb_log_destroy(ofb_log* log);

/* ---- primitive collider on the GPU ---------------------------------------------------
 * The contact-generation step before the solver (Newton's collide, called at
 * simulation/base_simulator.py:118 before integrator.simulate).  Newton's narrow phase is
 * a third-party dependency absent here; this collider generates contacts for the
 * benchmark primitives (plane, box, cylinder, sphere, heightfield) from fixed per-model
 * candidate points and writes them in the ofb_contacts format.  The description holds
 * HOST pointers (copied at creation); geometry codes are 0 plane, 1 box, 2 cylinder,
 * 3 sphere, 4 heightfield; per-shape arrays are [W*shape_cols] (xform 7 floats, size 3). */
typedef struct ofb_collider_desc {
  int32_t num_worlds;
  int32_t bodies_per_world;
  int32_t shape_cols;
  int32_t num_pairs;
  const int32_t* pairs;         /* [num_pairs*2] (column of shape0, column of plane/box) */
  const int32_t* shape_geo;     /* [W*SC] */
  const int32_t* shape_body;    /* [W*SC] world-local body, -1 static */
  const float* shape_xform;     /* [W*SC*7] */
  const float* shape_size;      /* [W*SC*3] */
  const int32_t* flat_shape_id; /* [W*SC] id written to shape0/shape1 */
  int32_t cylinder_samples;     /* k: 3k rim points of a cylinder against a box */
  float gap;                    /* candidate is a contact when its distance < gap */
  float lateral_slack;          /* against a box: allowed excess on the tangent axes */
  int32_t max_contacts_per_world;
  /* spheres on bodies against one heightfield column (num_hf_spheres = 0: none) */
  int32_t num_hf_spheres;
  const int32_t* hf_spheres;    /* [num_hf_spheres] shape columns */
  int32_t hf_col;
  int32_t hf_nx;
  int32_t hf_ny;
  float hf_origin[2];
  float hf_spacing[2];
  const float* hf_heights;      /* [hf_ny*hf_nx] row-major (y, x) */
  /* contacts per sphere: 1 = the bilinear surface below its centre; 4 = the four grid
   * cells around its nearest grid vertex, each at the closest point of the cell's nearer
   * triangle (cells outside the grid give none) */
  int32_t hf_points;
} ofb_collider_desc;

/* Output of ofb_collide: the same fields as ofb_contacts, writable (device pointers). */
typedef struct ofb_contact_buffers {
  int32_t* count; /* [1] */
  int32_t max_contacts;
  float* point0;
  float* point1;
  float* normal;
  int32_t* shape0;
  int32_t* shape1;
  float* margin0;
  float* margin1;
} ofb_contact_buffers;

typedef struct ofb_collider ofb_collider;
int ofb_collider_create(const ofb_collider_desc* desc, ofb_collider** out);
int ofb_collider_destroy(ofb_collider* collider);
/* Contacts of body poses body_q [W*B*7] (device), in (world, pair, point) order with the
 * heightfield spheres after the pairs, capped at W*max_contacts_per_world; count[0] is
 * written on the device (no host synchronisation, capturable).  Slots at and beyond
 * count[0] are left as they were. */
int ofb_collide(ofb_collider* collider, const float* body_q, const ofb_contact_buffers* out, void* stream);

/* Last error message of the calling thread ("" if none). */
const char* ofb_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* OSTRICH_FB_H */
