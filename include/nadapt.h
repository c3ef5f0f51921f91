/*
 * nadapt: MI355X (gfx950) batched step-doubling rigid-body step behind a C ABI.
 *
 * Every entry point takes plain device pointers and sizes; `stream` is a hipStream_t
 * passed as void* (NULL = default stream).  Status codes: 0 ok, 1 incomplete frame
 * (max_substeps reached), 2 inner convex solve failed, negative = error (message in
 * na_last_error_string()).  Nothing here allocates on the caller's behalf except
 * na_solver_create, which owns the controller vectors and scratch.
 *
 * Reference interfaces replaced (mardigiorgio/newton-adaptive, Python/Warp):
 *   na_solver_create   SolverSAPAdaptive.__init__        newton/_src/solvers/sap/solver_sap_adaptive.py:446-601
 *                      SolverMuJoCoAdaptive.__init__     newton/_src/solvers/mujoco/solver_mujoco_adaptive.py:443-580
 *   na_step            SolverSAPAdaptive.step/integrate  solver_sap_adaptive.py:910-975
 *                      SolverMuJoCoAdaptive.step         solver_mujoco_adaptive.py:774-849
 *   na_step_plain      the same boundary call as one batched launch per iteration plus the
 *                      4-byte flag read of _run_substep_loop (solver_sap_adaptive.py:877-907)
 *   na_reset           reset / _reset_worlds             solver_sap_adaptive.py:982-1001, solver_mujoco_adaptive.py:234,912
 *   na_collide         CollisionPipeline.collide         newton/_src/sim/collide.py:920-1250
 *   na_collide_debug   compute_shape_aabbs + broad phase collide.py:150-275, geometry/broad_phase_sap.py:510-677
 *   na_sort_contacts   make_contact_sort_key + ContactSorter.sort_full (deterministic=True)
 *                                                     geometry/contact_data.py:59-83, geometry/contact_sort.py:194-420,
 *                                                     sim/collide.py:1125-1144
 *   na_substep         SolverSAPAdaptive.substep         solver_sap_adaptive.py:714-729
 *   na_attempt         _substep_body (one step-doubling attempt, no commit) solver_sap_adaptive.py:732-849
 *   na_ctrl_*          controller kernels                solver_mujoco_adaptive.py:63-378, solver_sap_adaptive.py:237-345
 */
#ifndef NADAPT_H
#define NADAPT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model arrays (device pointers, SoA).  Layout: bodies of world w are
 * [world_body_start[w], world_body_start[w+1]); world-local shapes of world w are
 * [world_shape_start[w], world_shape_start[w+1]); world -1 shapes are listed in
 * global_shapes.  Mirrors newton Model fields (newton/_src/sim/model.py:36). */
typedef struct na_model_desc {
  int32_t world_count, body_count, shape_count, global_count, filter_count;
  const int32_t* world_body_start;  /* [W+1] */
  const int32_t* world_shape_start; /* [W+1] */
  const int32_t* global_shapes;     /* [G] */
  const float* body_com;            /* [B,3] */
  const float* body_mass;           /* [B] */
  const float* body_inv_mass;       /* [B] */
  const float* body_inertia;        /* [B,3,3] body frame */
  const float* body_inv_inertia;    /* [B,3,3] */
  const int32_t* shape_type;        /* GeoType codes */
  const int32_t* shape_body;
  const int32_t* shape_collision_group;
  const int32_t* shape_world;
  const float* shape_transform;     /* [S,7] p, q(xyzw) in body frame */
  const float* shape_scale;         /* [S,3] */
  const float* shape_margin;
  const float* shape_gap;
  const float* shape_material_ke;
  const float* shape_material_mu;
  const float* shape_collision_radius;
  const int32_t* filter_pairs;      /* [F,2] sorted canonical (min,max) */
  float gravity[3];
  /* Optional [W] arrays: one past each world's last body / shape.  NULL means the start arrays
     are packed prefix sums ([W+1], end[w] = start[w+1]); separate end arrays let a world pool
     hold worlds in arbitrary arena ranges with gaps (serving). */
  const int32_t* world_body_end;
  const int32_t* world_shape_end;
  /* Optional candidate pairs for the explicit broad phase (broad_phase 2; model.shape_contact_pairs
     of the reference): canonical (min, max) global shape indices passing the world, group and
     exclusion rules, grouped by world; world w's pairs are [world_pair_start[w],
     world_pair_start[w+1]).  explicit_pair_count 0 = none. */
  int32_t explicit_pair_count;
  const int32_t* explicit_pairs;
  const int32_t* world_pair_start;
  /* Optional articulation (joint_type NULL: every body is a free body).  One joint per body:
     joint j moves body j, parents come before children within one world, a world holds at most
     16 bodies, 32 coordinates and 32 dofs.  Newton's joint_q / joint_qd layout
     (state.py:149-153): REVOLUTE (type 1) one angle / rate; FREE (type 4, root only, parent -1)
     origin position + quaternion xyzw / COM linear velocity + angular velocity, world frame. */
  const int32_t* joint_type;       /* [B] */
  const int32_t* joint_parent;     /* [B] global parent body, -1 world */
  const int32_t* joint_q_start;    /* [B+1] */
  const int32_t* joint_qd_start;   /* [B+1] */
  const float* joint_X_p;          /* [B,7] joint frame in the parent body frame */
  const float* joint_X_c;          /* [B,7] joint frame in the child body frame */
  const float* joint_axis;         /* [B,3] revolute axis in the joint frame */
  const float* joint_target_pos;   /* [D] PD drive targets, gains and armature per dof */
  const float* joint_target_ke;
  const float* joint_target_kd;
  const float* joint_armature;
  /* Optional MuJoCo geom parameters per shape, needed by contact_law 1 only (the model.mujoco
     custom attributes of SolverMuJoCo, solver_mujoco.py:700-790, with solref resolved as
     solver_mujoco.py:5224-5233 does: an authored solref, else convert_solref(ke, kd, 1, 1)).
     solref [S,2] (time constant, damping ratio; <= 0: direct stiffness/damping), solimp [S,5]
     (dmin, dmax, width, midpoint, power), solmix [S], priority [S], condim [S] (1 or 3).
     mj_zero_margins: rows use the surface distance, as SolverMuJoCo zeroes the margins of models
     with boxes for its native collision (solver_mujoco.py:3352-3357); else dist - margins. */
  const float* mj_solref;
  const float* mj_solimp;
  const float* mj_solmix;
  const int32_t* mj_priority;
  const int32_t* mj_condim;
  int32_t mj_zero_margins;
} na_model_desc;

/* Solver settings (constructor kwargs of the reference solvers). */
typedef struct na_config {
  int32_t mode;            /* 0 "fixed", 1 "adaptive" */
  int32_t controller;      /* 0 SAP (_adapt_dt), 1 MuJoCo (_calc_adjusted_step) */
  double tol, dt_inner_init, dt_inner_min, dt_inner_max, divergence_threshold;
  int32_t max_substeps, max_iterations, max_rigid_contact, broad_phase; /* 0 sap, 1 nxn, 2 explicit */
  float optimality_rel_tol, optimality_abs_tol, contact_tau_d, sigma, beta, ls_rel_tol, cg_rel_tol;
  int32_t max_cg;
  float cg_forcing;        /* > 0: inexact-Newton PCG forcing bound (eta_k <= cg_forcing); 0: fixed cg_rel_tol */
  int32_t line_search;     /* 0 "exact_root" (root of l'(alpha), cubic seed), 1 "armijo_decay" (backtracking,
                              the reference default, solver_sap_adaptive.py:459) */
  int32_t contact_law;     /* 0 SAP compliant contact (SolverSAPAdaptive); 1 MuJoCo soft constraints with the
                              pyramidal cone (SolverMuJoCoAdaptive's inner step, mujoco_warp.step,
                              solver_mujoco_adaptive.py:585-605): needs the model's mj_* arrays; free bodies */
  float impratio;          /* MuJoCo frictional-to-normal impedance ratio (> 0; contact_law 1) */
  int32_t warm_start;      /* Newton seeds of an attempt's three solves.  0: the reference's
                              (solver_sap_adaptive.py:756-771): full step from v_t, first half step
                              from (v_t + v_full) / 2, second from v_full.  1: extrapolated: full step
                              from v_t + dt/2 a (a: the body's mean acceleration over its last
                              committed step, kept in body_trend), first half step as 0, second from
                              2 v_mid - v_t.  2: concurrent: as 1, but the first half step from
                              v_t + dt/4 a, so it does not depend on the full step's result.  Seeds
                              move where Newton starts, not what it converges to.  Free-body worlds;
                              articulated worlds use 0's seeds. */
} na_config;

/* Control inputs of a boundary call (newton Control, newton/_src/sim/control.py:31-110; applied by
 * the reference's SolverMuJoCo._apply_mjc_control, solvers/mujoco/solver_mujoco.py:3965-4085, and
 * sap_control_from_newton, solver_sap_adaptive.py:922).  Device float arrays in the model's
 * joint_qd (dof) layout; any pointer may be NULL.  joint_f: generalized forces -- a FREE joint's six
 * entries are a wrench (force; torque) at the child body's COM in the world frame, added to that
 * body's State.body_f; joint_target_q: position targets of the drives (dof layout, the reference's
 * default newton.use_coord_layout_targets = False), replacing the model's joint_target_pos;
 * joint_target_qd: velocity targets (NULL = 0); joint_act: feedforward generalized forces.
 * A drive applies tau = ke (q_t - q) + kd (qd_t - qd) + act, implicitly in the step. */
typedef struct na_control_desc {
  const float* joint_f;
  const float* joint_target_q;
  const float* joint_target_qd;
  const float* joint_act;
} na_control_desc;

/* Device pointers to the per-world controller vectors (owned by the solver). */
typedef struct na_telemetry {
  double* dt;
  double* dt_half;
  double* ideal_dt;
  double* sim_time;
  double* next_time;
  double* last_raw_error;
  double* last_error;          /* accepted error */
  uint8_t* accepted;
  uint8_t* commit;
  uint8_t* diverged;
  int32_t* substeps;           /* accepted steps in the last frame */
  int32_t* iterations;         /* attempts in the last frame */
  int32_t* solve_fail;
  int32_t* overflow;           /* contact/pair capacity overflows */
  unsigned long long* cum_accepted; /* [1] */
  long long* work;             /* [W,12]: attempts, newton iters, cg iters, contacts, pairs, solves,
                                  then cycles of collision, setup, residual, preconditioner, PCG,
                                  line search (zero unless built with -DNA_PHASE_TIMERS) */
  double* body_q64;            /* [B,7] FP64 pose accumulator */
  float* body_trend;           /* [B,6] mean acceleration of each body over its last committed step
                                  (warm_start 1; may be NULL otherwise); zero = no history */
} na_telemetry;

/* Contacts SoA output (newton/_src/sim/contacts.py:137-345). */
typedef struct na_contacts_desc {
  int32_t capacity;
  int32_t* count;              /* [1] total generated (may exceed capacity) */
  int32_t* shape0;
  int32_t* shape1;
  float* point0;               /* [C,3] body frame */
  float* point1;
  float* offset0;
  float* offset1;
  float* normal;               /* [C,3] world, A -> B */
  float* margin0;
  float* margin1;
  int32_t* world_count_out;    /* [W] contacts per world (before capacity clamp) */
  int32_t* scratch;            /* [W*per_world*16] staging words, see na_collide_scratch_words */
} na_contacts_desc;

typedef struct na_model na_model;
typedef struct na_solver na_solver;

const char* na_last_error_string(void);
int na_device_info(char* name, int name_len, int* cu_count, int* lds_per_block);

int na_model_create(const na_model_desc* desc, na_model** out);
int na_model_destroy(na_model* m);

/* `buffers`: caller-owned controller vectors (all pointers set, sized for the model), or
 * NULL to let the solver allocate them. */
int na_solver_create(const na_model* m, const na_config* cfg, const na_telemetry* buffers, na_solver** out);
int na_solver_destroy(na_solver* s);
int na_solver_telemetry(const na_solver* s, na_telemetry* out);
/* Schedule: persistent grid of `grid_waves` wavefronts (<=0: automatic) pulling worlds
 * in `order` (device [W] permutation, or NULL for natural order). */
int na_solver_set_schedule(na_solver* s, int grid_waves, const int32_t* order);
/* Two-wave tiles: the first `tiles` entries of a served launch's world list (na_serve) or of a
 * boundary call's visiting order (na_step; under the cost order, the previous call's costliest
 * worlds) each run on a workgroup of two wavefronts that solve an attempt's full step and first
 * half step at the same time; the rest runs one world per wavefront.  0 (default): no tiles.
 * Needs adaptive mode, warm_start 2 and contact_law 0; a tiled world's results equal its
 * single-wave results bit for bit.  Replaces the reference's single captured iteration body
 * for every world (solver_sap_adaptive.py:851-907) with a per-world schedule. */
int na_solver_set_tiles(na_solver* s, int32_t tiles);

/* Boundary call: advance every world by dt_outer in place (served form: one persistent
 * launch, on-device ragged loop).  body_f [B,6] (may be NULL): applied COM wrenches, world frame;
 * control (may be NULL): see na_control_desc (free bodies use joint_f only). */
int na_step(na_solver* s, float* body_q, float* body_qd, const float* body_f, const na_control_desc* control,
            double dt_outer, void* stream);
/* Same boundary call in the lock-step form: one launch over all worlds per iteration and a
 * 4-byte flag read back per iteration.  Returns the status of the host loop. */
/* Lock-step form's work pattern: with batch_all set, a world that has reached its boundary still
 * runs each iteration's two collision passes and discards them, as the reference's batched
 * CollisionPipeline.collide does for every world of its fixed batch (solver_sap_adaptive.py:754,769;
 * its zero-length solves are not charged).  Results are the same either way; 0 (default) skips
 * finished worlds. */
int na_solver_set_plain_batch(na_solver* s, int32_t batch_all);
int na_step_plain(na_solver* s, float* body_q, float* body_qd, const float* body_f, const na_control_desc* control,
                  double dt_outer, int32_t* iterations_out, void* stream);
/* Continuous-batching rollout launch over the worlds listed in `active` (no counterpart in the
   reference: the serving form of the boundary call).  Each listed world w runs boundary frames of
   dt_outer until frames_done[w] reaches horizon_frames[w] or accepted[w] reaches
   horizon_accepted[w] (status[w] = 1), its convex solve fails (status 2), or budget_ms of device
   time has passed since the launch began (status stays 0; the world resumes exactly in a later
   call).  A rollout of F frames leaves the same state as F calls of na_step.  world_tol (may be
   NULL) gives a per-world error tolerance.  max_iterations > 0 caps the step-doubling iterations
   each world takes in this call (1 = the lock-step form of the reference's host loop);
   unfinished (device int32, may be NULL) is set to 1 when any listed world is still running.
   Asynchronous on `stream`; budget_ms <= 0 means no budget. */
int na_serve(na_solver* s, const int32_t* active, int32_t n_active, const int32_t* horizon_frames,
             const int32_t* horizon_accepted, int32_t* frames_done, int32_t* accepted, int32_t* status,
             const double* world_tol, float* body_q, float* body_qd, const float* body_f, double dt_outer,
             double budget_ms, int32_t max_iterations, int32_t* unfinished, void* stream);

/* MuJoCo coordinate adapters for the FREE joint every body carries (parent frame = world).
   Newton joint_q [B,7] = origin position + quaternion xyzw, joint_qd [B,6] = COM linear velocity
   + angular velocity (world); MuJoCo qpos [B,7] = origin position + quaternion wxyz, qvel [B,6] =
   origin linear velocity (world) + angular velocity (body frame).  Replace
   convert_warp_coords_to_mj_kernel (solvers/mujoco/kernels.py:705-789) and
   convert_mj_coords_to_warp_kernel (kernels.py:578-666) for FREE joints.  Rows of one world are
   contiguous, so equal-size worlds give MuJoCo's [nworld, nq] layout.
   Articulated models: qpos / qvel mirror the joint_q / joint_qd layout; a FREE root converts as
   above with its twist (COM velocity, angular velocity) taken in the joint's parent frame
   (joint_X_p), revolute coordinates and velocities are copied (no reference offset). */
int na_coords_to_mujoco(const na_model* m, const float* joint_q, const float* joint_qd, float* qpos, float* qvel,
                        void* stream);
int na_coords_from_mujoco(const na_model* m, const float* qpos, const float* qvel, float* joint_q, float* joint_qd,
                          void* stream);

/* Frame-to-frame contact matching (replaces ContactMatcher.match, geometry/contact_match.py:566-,
   called from collide.py:853-861 when contact_matching is on).  `cur` is the buffer na_collide just
   filled (cur_offsets: its per-world start offsets, the W words after the staging area of the
   na_collide scratch); prev_* describe the previous frame's buffer (shape pairs, world-space
   midpoints from the previous call's cur_mid, normals, per-world offsets and unclamped counts).
   Writes match_index for every current contact (>= 0 previous index, -1 new pair, -2 broken),
   the current world midpoints into cur_mid [capacity,3] and prev_matched[k] = 1 for every
   previous contact that was matched (prev_matched must be zero on entry).  Sticky mode
   (contact_matching="sticky", ContactMatcher.replay_matched, contact_match.py; all four prev_point0/1,
   prev_offset0/1 [prev_capacity,3] set, else all NULL): a matched contact takes the previous
   contact's body-frame points, offsets and normal, and its midpoint is recomputed from them. */
int na_match_contacts(const na_model* m, const float* body_q, const na_contacts_desc* cur, int32_t per_world,
                      const int32_t* cur_offsets, const int32_t* prev_shape0, const int32_t* prev_shape1,
                      const float* prev_mid, const float* prev_normal, const int32_t* prev_offsets,
                      const int32_t* prev_counts, int32_t prev_capacity, float pos_threshold,
                      float normal_dot_threshold, int32_t* match_index, float* cur_mid, int32_t* prev_matched,
                      const float* prev_point0, const float* prev_point1, const float* prev_offset0,
                      const float* prev_offset1, void* stream);

/* Deterministic contact order (replaces make_contact_sort_key, geometry/contact_data.py:59-83, and
   ContactSorter.sort_full, geometry/contact_sort.py:194-420, run by collide.py:1125-1144 when the
   pipeline is deterministic).  Sorts the first min(*c->count, c->capacity) contacts of `c` in place
   by the 63-bit key  shape0 << 43 | shape1 << 23 | sub  (20/20/23-bit fields; sub = rank of the
   contact within its shape pair in production order), over the whole capacity with the sentinel
   0x7FFFFFFFFFFFFFFF past the count, on `stream` without host synchronization.
   keys_out [capacity] int64 receives the sorted keys.  Optional (NULL to skip): tids [capacity] and
   match_index [capacity] are permuted with the contacts; a match index k >= 0 is replaced by
   match_remap[k] (k < match_remap_len), which maps the previous frame's pre-sort positions to its
   sorted ones; rank_out [capacity] receives, for every pre-sort position, its sorted position.
   shape_count bounds the shape ids (<= 2^20 - 1).  workspace: device memory of at least
   na_sort_contacts_workspace_bytes(capacity) bytes (returns -1 on error). */
int64_t na_sort_contacts_workspace_bytes(int32_t capacity);
int na_sort_contacts(const na_contacts_desc* c, int32_t shape_count, int32_t* tids, int32_t* match_index,
                     const int32_t* match_remap, int32_t match_remap_len, int64_t* keys_out, int32_t* rank_out,
                     void* workspace, int64_t workspace_bytes, void* stream);

/* Masked controller reset (world_mask: device uint8 [W], NULL = all worlds); also clears the
   reset worlds' body_trend. */
int na_reset(na_solver* s, const uint8_t* world_mask, void* stream);

/* Contact generation into the Contacts arrays, world-major deterministic order. */
int64_t na_collide_scratch_words(const na_model* m, int32_t per_world);
int na_collide(const na_model* m, const float* body_q, const na_contacts_desc* out, int32_t per_world,
               int32_t broad_phase, void* stream);
/* Allocation-free variant of na_collide.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels and outputs as na_collide (out->scratch is ignored); the staging
 * area and per-world offsets live in a static device array of NA_COLLIDE_NOALLOC_MAX_WORDS int32
 * words (48 MiB, held by every GPU the library runs on until the process ends).  Requires
 * world_count * per_world * 20 + 2 * world_count <= NA_COLLIDE_NOALLOC_MAX_WORDS, else returns -1.
 * Calls on one GPU share that array, so calls that overlap on one GPU (other streams, other host
 * threads) race; the caller orders them. */
#define NA_COLLIDE_NOALLOC_MAX_WORDS (12 * 1024 * 1024)
int na_collide_noalloc(const na_model* m, const float* body_q, const na_contacts_desc* out, int32_t per_world,
                       int32_t broad_phase, void* stream);
/* AABBs [S,3] lo/hi and per-world candidate pairs [W, max_pairs, 2] (global shape ids,
 * canonical) with counts [W]. */
int na_collide_debug(const na_model* m, const float* body_q, float* aabb_lo, float* aabb_hi, int32_t* pairs,
                     int32_t* pair_count, int32_t max_pairs, int32_t broad_phase, void* stream);

/* One inner step per world (contacts at q, convex solve at h[w], FP64 integration).
 * guess [B,6] FP64 (NULL: start at qd) seeds the Newton iterate, as `substep(..., guess)` does
 * (solver_sap_adaptive.py:714-729); a seed that already meets the optimality test takes zero
 * Newton iterations.  v_solution [B,6] FP64 (NULL to skip) receives the converged velocity.
 * stats [W,6] int32: Newton iterations, CG iterations, converged, contacts, exit reason
 * (0 converged, 1 iteration cap, 2 no descent, 3 preconditioner not SPD, 4 NaN), final
 * relative residual (float bits). */
int na_substep(const na_solver* s, const double* q64, const float* qd, const float* f, const double* h,
               const double* guess, double* q64_out, float* qd_out, double* v_solution, int32_t* stats,
               void* stream);
/* One step-doubling attempt per world from (q64, qd) with dt[w]: full and doubled results,
 * FP64 error norm, solve status and contact count; no commit, no controller update.
 * contacts_out [W, 2, 128, 16] float and contact_counts_out [W, 2] int32 (both NULL to skip): the
 * contact sets the attempt solved with, at q and at the midpoint state, per contact (shape a,
 * shape b, body a, body b (global, -1 static), manifold slot, unit normal A->B, world midpoint,
 * distance, phi0 = distance - margins, stiffness, friction, 0) -- for parity checks that give a
 * reference solve the same contacts.  trend [B,6] (may be NULL = zero): the bodies' accelerations
 * for the full step's seed under warm_start 1 (ignored under 0). */
int na_attempt(const na_solver* s, const double* q64, const float* qd, const float* f, const double* dt,
               double* q_full, float* qd_full, double* q_dbl, float* qd_dbl, double* err, int32_t* solve_ok,
               int32_t* n_contacts, float* contacts_out, int32_t* contact_counts_out, const float* trend,
               void* stream);

/* Articulated models (model created with joint arrays).  The boundary call on joint coordinates:
   every world advances by dt_outer with step doubling over its joint_q / joint_qd (error =
   max-norm of the coordinate difference, _inf_norm_state_error_kernel, solver_sap_adaptive.py:
   154-185; attempt sequence _substep_body, :732-849); replaces SolverSAPAdaptive.step (:962) for
   models with joints.  joint_q / joint_qd are updated in place (an FP64 coordinate shadow is kept
   while it still rounds to joint_q); body_q / body_qd (may be NULL) receive the bodies' poses and
   COM / angular velocities at the new coordinates (eval_fk).  body_f [B,6] (may be NULL): applied
   COM wrenches, world frame; control (may be NULL): joint forces, drive targets and feedforward
   (na_control_desc), held for the whole call. */
int na_step_articulated(na_solver* s, float* joint_q, float* joint_qd, float* body_q, float* body_qd,
                        const float* body_f, const na_control_desc* control, double dt_outer, void* stream);
/* Serving form of na_step_articulated (no counterpart in the reference, like na_serve): the
   articulated worlds listed in `active` run boundary frames until their horizon, a solve failure
   or the device-time budget, with the semantics of na_serve.  joint_q64 [coordinates] and
   joint_qd64 [dofs] are caller-owned FP64 shadows: a world resumes from them while they still
   round to joint_q / joint_qd, so budget interruptions are exact; velocities are rounded to
   float32 at every frame boundary, so F frames equal F calls of na_step_articulated. */
int na_serve_articulated(na_solver* s, const int32_t* active, int32_t n_active, const int32_t* horizon_frames,
                         const int32_t* horizon_accepted, int32_t* frames_done, int32_t* accepted, int32_t* status,
                         const double* world_tol, float* joint_q, float* joint_qd, double* joint_q64,
                         double* joint_qd64, float* body_q, float* body_qd, double dt_outer, double budget_ms,
                         int32_t max_iterations, int32_t* unfinished, void* stream);
/* Serving launch over a pool holding free-body and articulated worlds (no counterpart in the
   reference): one persistent launch in which every wave takes the next listed world of either kind,
   so both kinds share the device; per world the arithmetic of na_serve (body_q / body_qd / body_f)
   or na_serve_articulated (joint_q / joint_qd with their FP64 shadows, body_q / body_qd by forward
   kinematics).  A world is articulated when its first body has joint_type != 0.  body_f (may be
   NULL) applies to the bodies of both kinds. */
int na_serve_mixed(na_solver* s, const int32_t* active, int32_t n_active, const int32_t* horizon_frames,
                   const int32_t* horizon_accepted, int32_t* frames_done, int32_t* accepted, int32_t* status,
                   const double* world_tol, float* body_q, float* body_qd, const float* body_f, float* joint_q,
                   float* joint_qd, double* joint_q64, double* joint_qd64, double dt_outer, double budget_ms,
                   int32_t max_iterations, int32_t* unfinished, void* stream);
/* One inner step per articulated world (contacts at q, convex solve at h[w], FP64 integration);
   the parity form of substep (solver_sap_adaptive.py:714-729).  body_f / control as in
   na_step_articulated (may be NULL).  stats [W,4] int32: Newton iterations, factorizations,
   converged, contacts. */
int na_art_substep(const na_solver* s, const double* joint_q, const double* joint_qd, const double* h,
                   const float* body_f, const na_control_desc* control, double* q_out, double* qd_out,
                   int32_t* stats, void* stream);
/* One step-doubling attempt per articulated world from (joint_q, joint_qd) with dt[w]: full and
   doubled coordinates / velocities, FP64 error norm, solve status, contacts at q; contacts_out /
   contact_counts_out as in na_attempt (may be NULL). */
int na_art_attempt(const na_solver* s, const double* joint_q, const double* joint_qd, const double* dt,
                   const float* body_f, const na_control_desc* control, double* q_full, double* qd_full,
                   double* q_dbl, double* qd_dbl, double* err, int32_t* solve_ok, int32_t* n_contacts,
                   float* contacts_out, int32_t* contact_counts_out, void* stream);

/* Controller kernels with the reference float32 arithmetic (one thread per world). */
int na_ctrl_calc_adjusted_step_f32(int32_t n, const float* err, const float* dt, float* ideal_dt, uint8_t* accepted,
                                   uint8_t* commit, uint8_t* diverged, float tol, float dt_min, float div_thr,
                                   void* stream);
int na_ctrl_adapt_dt_f32(int32_t n, const float* err, float* sim_time, const float* next_time, float* dt,
                         float* dt_half, float* ideal_dt, uint8_t* diverged, uint8_t* accept, float* accepted_error,
                         int32_t* substeps, int32_t* cum_accepted, int32_t mode, float tol, float dt_min,
                         float dt_max, float div_thr, void* stream);
int na_ctrl_rebase_time_f32(int32_t n, float* sim_time, float* next_time, void* stream);
int na_ctrl_reset_worlds_f32(int32_t n, const uint8_t* mask, float dt_init, float* ideal_dt, float* dt,
                             float* dt_half, float* sim_time, float* next_time, uint8_t* diverged,
                             uint8_t* accepted, void* stream);
int na_ctrl_clamp_dt_to_boundary_f32(int32_t n, float* dt, float* dt_half, const float* sim_time,
                                     const float* next_time, void* stream);
int na_ctrl_inf_norm_error_f32(int32_t n, int32_t coords, const float* q_full, const float* q_double,
                               const float* scale, float* err, void* stream);

#ifdef __cplusplus
}
#endif
#endif
