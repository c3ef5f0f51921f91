/*
 * CPU oracle for the batched step-doubling rigid-body step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path never links or calls it.
 *
 * It restates, in plain C over REAL (double for the oracle build, float for the
 * CPU-baseline port), the reference algorithm of mardigiorgio/newton-adaptive:
 *
 *   controller   newton/_src/solvers/mujoco/solver_mujoco_adaptive.py:63-415
 *                newton/_src/solvers/sap/solver_sap_adaptive.py:89-434,732-849,877-960
 *   AABBs        newton/_src/sim/collide.py:150-275
 *   broad phase  newton/_src/geometry/broad_phase_common.py:18-180 (brute-force n^2 set,
 *                the semantics of newton/tests/test_broad_phase.py:90-148)
 *   narrow phase newton/_src/geometry/narrow_phase.py:154-604,
 *                newton/_src/geometry/collision_primitive.py:49-1215
 *   contacts     newton/_src/sim/collide.py:62-147, geometry/contact_data.py:59-117
 *   integrator   newton/_src/solvers/solver.py:50-94
 *   inner solve  the convex SAP problem the reference solves through the absent
 *                external sap_warp (call sites solver_sap_adaptive.py:489-501,714-729),
 *                restated from the published SAP formulation (Castro et al. 2022) and
 *                CENIC (Kurtz & Castro, arXiv:2511.08771): primal Newton on
 *                l(v) = 1/2|v-v*|_A^2 + sum_i 1/2 |P_F(y_i)|_R^2, dense Cholesky,
 *                exact line search seeded by the cubic Hermite minimiser (N&W 3.59).
 *
 * Box-box and capsule-box, which the reference routes to GJK/MPR + polygon
 * clipping (narrow_phase.py:629-779), are restated here by this project's own
 * SAT + face clipping and segment/box signed-distance minimisation (the same
 * algorithm the HIP path implements).
 *
 * Pinning (see tests/): controller KATs from test_adaptive_floor_nan_guard.py,
 * test_adaptive_time_rebase.py, test_adaptive_reset_worlds.py; cubic KATs from
 * test_adaptive_cubic_linesearch.py; primitive closed-form tables from
 * test_collision_primitives.py; brute-force pair sets; solver optimality and
 * invariants (free fall closed form, momentum, energy bound, rest); articulated
 * mechanics identities (host FK, kinetic energy, Coriolis/gravity power, floating base
 * equal to the free body, PD torque balance, dt^2 error order).
 *
 * Parity unpinned (no reference values can be produced here, the dependencies are
 * absent): the exact manifold points of the reference's box-box / capsule-box MPR
 * multicontact, the values of the reference's sap_warp convex solve, and the
 * mujoco_warp inner step behind SolverMuJoCoAdaptive.  Those parts are checked by the
 * invariants above only.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef OR_FLOAT
typedef float real;
#define RS(x) sqrtf(x)
#define RA(x) fabsf(x)
#else
typedef double real;
#define RS(x) sqrt(x)
#define RA(x) fabs(x)
#endif

/* `acc` carries what the device path keeps in FP64 whatever `real` is: the Newton iterate, the
   gradient, the cost, the friction-cone projection and the line search.  The FP32 build (OR_FLOAT,
   the CPU baseline) thus runs the device's precision placement -- FP32 geometry, contact rows,
   Hessian and factor, FP64 residual -- and the FP64 build is unchanged. */
typedef double acc;

#define OR_MAXVAL 1.0e10
#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------ vectors */
typedef struct { real x, y, z; } v3;
typedef struct { real x, y, z, w; } qt;

static inline v3 V(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, real s) { return V(a.x * s, a.y * s, a.z * s); }
static inline real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline real len(v3 a) { return RS(dot(a, a)); }
static inline real comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void setc(v3* a, int i, real s) { if (i == 0) a->x = s; else if (i == 1) a->y = s; else a->z = s; }
static inline v3 nrm(v3 a) { real l = len(a); return l > 0 ? scl(a, (real)1 / l) : a; }
static inline real clampr(real x, real lo, real hi) { return x < lo ? lo : (x > hi ? hi : x); }
static inline real maxr(real a, real b) { return a > b ? a : b; }
static inline real minr(real a, real b) { return a < b ? a : b; }

/* rotation matrix (row-major) of quaternion (x, y, z, w) */
typedef struct { real m[9]; } m33;
static m33 qmat(qt q) {
  m33 r;
  real x = q.x, y = q.y, z = q.z, w = q.w;
  r.m[0] = 1 - 2 * (y * y + z * z); r.m[1] = 2 * (x * y - z * w);     r.m[2] = 2 * (x * z + y * w);
  r.m[3] = 2 * (x * y + z * w);     r.m[4] = 1 - 2 * (x * x + z * z); r.m[5] = 2 * (y * z - x * w);
  r.m[6] = 2 * (x * z - y * w);     r.m[7] = 2 * (y * z + x * w);     r.m[8] = 1 - 2 * (x * x + y * y);
  return r;
}
static inline v3 mv(const m33* a, v3 v) {
  return V(a->m[0] * v.x + a->m[1] * v.y + a->m[2] * v.z, a->m[3] * v.x + a->m[4] * v.y + a->m[5] * v.z,
           a->m[6] * v.x + a->m[7] * v.y + a->m[8] * v.z);
}
static inline v3 mtv(const m33* a, v3 v) {
  return V(a->m[0] * v.x + a->m[3] * v.y + a->m[6] * v.z, a->m[1] * v.x + a->m[4] * v.y + a->m[7] * v.z,
           a->m[2] * v.x + a->m[5] * v.y + a->m[8] * v.z);
}
static inline v3 col(const m33* a, int j) { return V(a->m[j], a->m[3 + j], a->m[6 + j]); }
static qt qmul(qt a, qt b) {
  qt r;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
  r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  return r;
}
static inline v3 qrot(qt q, v3 v) { m33 r = qmat(q); return mv(&r, v); }

/* ------------------------------------------------------------------ model */
typedef struct or_model {
  int world_count, body_count, shape_count, global_count;
  const int32_t *world_body_start, *world_shape_start, *global_shapes;
  const double *body_com, *body_mass, *body_inv_mass, *body_inertia, *body_inv_inertia;
  const int32_t *shape_type, *shape_body, *shape_group, *shape_world;
  const double *shape_xform, *shape_scale, *shape_margin, *shape_gap, *shape_ke, *shape_mu, *shape_radius;
  int filter_count;
  const int32_t* filter_pairs; /* sorted canonical (min, max) */
  double gravity[3];
  /* MuJoCo geom parameters per shape (solver_mujoco.py:700-790 custom attributes, resolved):
     solref [S,2] (time constant, damping ratio; or direct stiffness/damping when <= 0), solimp
     [S,5], solmix [S], priority [S], condim [S]; zero_margins: the model's margins are zeroed as
     SolverMuJoCo does for models with boxes (solver_mujoco.py:3352-3357).  NULL when unused. */
  const double *mj_solref, *mj_solimp, *mj_solmix;
  const int32_t *mj_priority, *mj_condim;
  int mj_zero_margins;
} or_model;

typedef struct or_config {
  int mode;            /* 0 fixed, 1 adaptive (solver_sap_adaptive.py:64-65) */
  int controller;      /* 0 SAP _adapt_dt, 1 MuJoCo _calc_adjusted_step */
  double tol, dt_init, dt_min, dt_max, divergence_threshold;
  int max_substeps, max_iterations, max_contacts;
  double opt_rel_tol, opt_abs_tol, tau_d, sigma, beta, ls_rel_tol;
  int line_search;     /* 0 exact_root, 1 armijo_decay (solver_sap_adaptive.py:459) */
  int contact_law;     /* 0 SAP compliant cone, 1 MuJoCo soft constraints, pyramidal cone */
  double impratio;     /* MuJoCo frictional-to-normal impedance ratio (law 1) */
  int warm_start;      /* Newton seeds of an attempt: 0 the reference's, 1 extrapolated, 2 concurrent (see attempt) */
} or_config;

enum { GEO_PLANE = 1, GEO_SPHERE = 3, GEO_CAPSULE = 4, GEO_BOX = 7 };

/* ====================================================================
 * Controller, float32 semantics exactly as the reference kernels.
 * ==================================================================== */
#define DRAKE_SAFETY 0.9
#define DRAKE_MIN_SHRINK 0.1
#define DRAKE_MAX_GROW 5.0
#define DRAKE_HYST_HIGH 1.2
#define DRAKE_HYST_LOW 0.9

/* _calc_adjusted_step, solver_mujoco_adaptive.py:120-198 */
#define DEFINE_CALC_ADJUSTED(T, SFX, SQ)                                                                   \
  void or_calc_adjusted_step_##SFX(int n, const T* err, const T* dt, T* ideal_dt, uint8_t* accepted,      \
                                   uint8_t* commit, uint8_t* diverged, T tol, T dt_min, T div_thr) {       \
    for (int w = 0; w < n; ++w) {                                                                          \
      T e = err[w], step = dt[w];                                                                          \
      int is_div = isnan(e) || isinf(e) || e >= div_thr;                                                  \
      if (step <= (T)0) { accepted[w] = 1; commit[w] = 1; continue; }                                     \
      if (step <= dt_min * (T)1.001) {                                                                     \
        if (is_div) { accepted[w] = 1; commit[w] = 0; diverged[w] = 1; ideal_dt[w] = dt_min; continue; }  \
        if (e > tol) { accepted[w] = 1; commit[w] = 1; ideal_dt[w] = dt_min; continue; }                  \
      }                                                                                                    \
      if (is_div) { accepted[w] = 0; commit[w] = 0; ideal_dt[w] = (T)DRAKE_MIN_SHRINK * step; continue; } \
      T ee = e > (T)1.0e-30 ? e : (T)1.0e-30;                                                              \
      T ns = (T)DRAKE_SAFETY * step * SQ(tol / ee);                                                        \
      if (ns > (T)DRAKE_HYST_LOW * step && ns < (T)DRAKE_HYST_HIGH * step) ns = step;                     \
      T lo = (T)DRAKE_MIN_SHRINK * step, hi = (T)DRAKE_MAX_GROW * step;                                    \
      ns = ns < lo ? lo : (ns > hi ? hi : ns);                                                             \
      int acc = (e <= tol) || (ns >= step);                                                                \
      accepted[w] = (uint8_t)acc; commit[w] = (uint8_t)acc; ideal_dt[w] = ns;                              \
    }                                                                                                      \
  }
DEFINE_CALC_ADJUSTED(float, f32, sqrtf)
DEFINE_CALC_ADJUSTED(double, f64, sqrt)

/* _adapt_dt, solver_sap_adaptive.py:237-345 */
#define DEFINE_ADAPT_DT(T, SFX, SQ)                                                                         \
  void or_adapt_dt_##SFX(int n, const T* err, T* sim_time, const T* next_time, T* dt, T* dt_half,          \
                         T* ideal_dt, uint8_t* diverged, uint8_t* accept, T* accepted_error,              \
                         int32_t* substeps, int32_t* cum_accepted, int mode, T tol, T dt_min, T dt_max,    \
                         T div_thr) {                                                                      \
    for (int w = 0; w < n; ++w) {                                                                          \
      T step = dt[w];                                                                                      \
      if (sim_time[w] >= next_time[w] || step <= (T)0) { accept[w] = 0; continue; }                       \
      T e = err[w];                                                                                        \
      int is_div = isnan(e) || isinf(e) || e >= div_thr;                                                  \
      if (mode == 0) {                                                                                     \
        if (is_div) { accept[w] = 0; sim_time[w] = next_time[w]; diverged[w] = 1; continue; }            \
        accept[w] = 1; sim_time[w] = sim_time[w] + step; accepted_error[w] = e; substeps[w] += 1;          \
        cum_accepted[0] += 1; continue;                                                                    \
      }                                                                                                    \
      T ee = e > (T)1.0e-30 ? e : (T)1.0e-30;                                                              \
      if (step <= dt_min * (T)1.001) {                                                                     \
        if (is_div) { accept[w] = 0; sim_time[w] = next_time[w]; diverged[w] = 1; ideal_dt[w] = dt_min;   \
                      continue; }                                                                          \
        accept[w] = 1; sim_time[w] = sim_time[w] + step; accepted_error[w] = e; substeps[w] += 1;          \
        cum_accepted[0] += 1;                                                                              \
        T ns = (T)DRAKE_SAFETY * step * SQ(tol / ee);                                                      \
        if (ns > (T)DRAKE_HYST_LOW * step && ns < (T)DRAKE_HYST_HIGH * step) ns = step;                   \
        T lo = (T)DRAKE_MIN_SHRINK * step, hi = (T)DRAKE_MAX_GROW * step;                                  \
        ideal_dt[w] = ns < lo ? lo : (ns > hi ? hi : ns);                                                  \
        continue;                                                                                          \
      }                                                                                                    \
      if (is_div) {                                                                                        \
        accept[w] = 0; T ns = (T)DRAKE_MIN_SHRINK * step; ideal_dt[w] = ns;                                \
        T d = ns < dt_min ? dt_min : (ns > dt_max ? dt_max : ns); dt[w] = d; dt_half[w] = d * (T)0.5;     \
        continue;                                                                                          \
      }                                                                                                    \
      T ns = (T)DRAKE_SAFETY * step * SQ(tol / ee);                                                        \
      if (ns > (T)DRAKE_HYST_LOW * step && ns < (T)DRAKE_HYST_HIGH * step) ns = step;                     \
      T lo = (T)DRAKE_MIN_SHRINK * step, hi = (T)DRAKE_MAX_GROW * step;                                    \
      ns = ns < lo ? lo : (ns > hi ? hi : ns);                                                             \
      int acc = (e <= tol) || (ns >= step);                                                                \
      T d = ns < dt_min ? dt_min : (ns > dt_max ? dt_max : ns);                                            \
      if (acc) { accept[w] = 1; sim_time[w] = sim_time[w] + step; accepted_error[w] = e; substeps[w] += 1; \
                 cum_accepted[0] += 1; }                                                                   \
      else accept[w] = 0;                                                                                  \
      ideal_dt[w] = ns; dt[w] = d; dt_half[w] = d * (T)0.5;                                                \
    }                                                                                                      \
  }
DEFINE_ADAPT_DT(float, f32, sqrtf)
DEFINE_ADAPT_DT(double, f64, sqrt)

/* _rebase_time, solver_mujoco_adaptive.py:340-358 */
void or_rebase_time_f32(int n, float* sim_time, float* next_time) {
  for (int i = 0; i < n; ++i) { float b = next_time[i]; sim_time[i] = sim_time[i] - b; next_time[i] = next_time[i] - b; }
}
/* _reset_worlds, solver_mujoco_adaptive.py:234-261 */
void or_reset_worlds_f32(int n, const uint8_t* mask, float dt_init, float* ideal_dt, float* dt, float* dt_half,
                         float* sim_time, float* next_time, uint8_t* diverged, uint8_t* accepted) {
  for (int i = 0; i < n; ++i) {
    if (!mask[i]) continue;
    ideal_dt[i] = dt_init; dt[i] = dt_init; dt_half[i] = dt_init * 0.5f;
    sim_time[i] = 0.0f; next_time[i] = 0.0f; diverged[i] = 0; accepted[i] = 0;
  }
}
/* _clamp_dt_to_boundary, solver_mujoco_adaptive.py:361-378 */
void or_clamp_dt_to_boundary_f32(int n, float* dt, float* dt_half, const float* sim_time, const float* next_time) {
  for (int i = 0; i < n; ++i) {
    float rem = next_time[i] - sim_time[i];
    if (rem <= 0.0f) { dt[i] = 0.0f; dt_half[i] = 0.0f; }
    else if (dt[i] > rem) { dt[i] = rem; dt_half[i] = rem * 0.5f; }
  }
}
/* _inf_norm_state_error_kernel, solver_mujoco_adaptive.py:78-108 */
void or_inf_norm_error_f32(int n, int coords, const float* q_full, const float* q_double, const float* scale,
                           float* err) {
  for (int w = 0; w < n; ++w) {
    float m = 0.0f;
    int nan_seen = 0;
    for (int i = 0; i < coords; ++i) {
      float d = fabsf(q_double[w * coords + i] - q_full[w * coords + i]);
      float s = scale[w * coords + i] * d;
      if (isnan(s)) nan_seen = 1; /* a non-finite coordinate poisons the world's error */
      m = m > s ? m : s;
    }
    if (nan_seen || isnan(m) || isinf(m)) m = 1.0e10f;
    err[w] = m;
  }
}

/* cubic Hermite minimiser on [0, h] (N&W eq. 3.59), test_adaptive_cubic_linesearch.py:36-60.
 * Returns 1 and *alpha on success, 0 on the fallback conditions. */
int or_cubic_min(double c0, double d0, double c1, double d1, double h, double* alpha) {
  double theta = d0 + d1 - 3.0 * (c1 - c0) / h;
  double disc = theta * theta - d0 * d1;
  if (!isfinite(disc) || disc < 0.0) return 0;
  double gamma = sqrt(disc);
  double denom = d1 - d0 + 2.0 * gamma;
  if (denom == 0.0) return 0;
  double a = h - h * (d1 + gamma - theta) / denom;
  if (!isfinite(a) || a <= 0.0 || a > h) return 0;
  *alpha = a;
  return 1;
}

/* ====================================================================
 * Narrow-phase primitives (collision_primitive.py).  Conventions:
 * normal from A into B, position = midpoint between surfaces,
 * distance negative = penetration.
 * ==================================================================== */
/* up to 8 contacts: box-box keeps its whole clipped face polygon */
#define OR_MAXMF 8
typedef struct { int n; real dist[OR_MAXMF]; v3 pos[OR_MAXMF]; v3 normal[OR_MAXMF]; } manifold;

static void mf_init(manifold* m) { m->n = 0; for (int i = 0; i < OR_MAXMF; ++i) m->dist[i] = (real)OR_MAXVAL; }

/* closest_segment_point, collision_primitive.py:49-54 */
static v3 closest_segment_point(v3 a, v3 b, v3 pt) {
  v3 ab = sub(b, a);
  real t = dot(sub(pt, a), ab) / (dot(ab, ab) + (real)1e-6);
  return add(a, scl(ab, clampr(t, 0, 1)));
}

/* collide_plane_sphere, :101-107 */
static void collide_plane_sphere(v3 n, v3 ppos, v3 spos, real r, real* dist, v3* pos) {
  real d = dot(sub(spos, ppos), n) - r;
  *dist = d;
  *pos = sub(spos, scl(n, r + (real)0.5 * d));
}

/* collide_sphere_sphere, :111-140 */
static void collide_sphere_sphere(v3 p1, real r1, v3 p2, real r2, real* dist, v3* pos, v3* n) {
  v3 dir = sub(p2, p1);
  real d = len(dir);
  v3 nn = d == 0 ? V(1, 0, 0) : scl(dir, (real)1 / d);
  d = d - (r1 + r2);
  *dist = d;
  *pos = add(p1, scl(nn, r1 + (real)0.5 * d));
  *n = nn;
}

/* collide_sphere_capsule, :144-176 */
static void collide_sphere_capsule(v3 sp, real sr, v3 cp, v3 axis, real cr, real hl, real* dist, v3* pos, v3* n) {
  v3 seg = scl(axis, hl);
  v3 pt = closest_segment_point(sub(cp, seg), add(cp, seg), sp);
  collide_sphere_sphere(sp, sr, pt, cr, dist, pos, n);
}

/* collide_capsule_capsule, :181-274 */
static void collide_capsule_capsule(v3 p1, v3 a1, real r1, real h1, v3 p2, v3 a2, real r2, real h2, manifold* m) {
  v3 axis1 = scl(a1, h1), axis2 = scl(a2, h2), dif = sub(p1, p2);
  real ma = dot(axis1, axis1), mb = -dot(axis1, axis2), mc = dot(axis2, axis2);
  real u = -dot(axis1, dif), v = dot(axis2, dif);
  real det = ma * mc - mb * mb;
  v3 nrm0;
  if (RA(det) >= (real)1e-15) {
    real inv = (real)1 / det;
    real x1 = (mc * u - mb * v) * inv, x2 = (ma * v - mb * u) * inv;
    if (x1 > 1) { x1 = 1; x2 = (v - mb) / mc; }
    else if (x1 < -1) { x1 = -1; x2 = (v + mb) / mc; }
    if (x2 > 1) { x2 = 1; x1 = clampr((u - mb) / ma, -1, 1); }
    else if (x2 < -1) { x2 = -1; x1 = clampr((u + mb) / ma, -1, 1); }
    v3 v1 = add(p1, scl(axis1, x1)), v2 = add(p2, scl(axis2, x2));
    collide_sphere_sphere(v1, r1, v2, r2, &m->dist[0], &m->pos[0], &nrm0);
    m->normal[0] = nrm0;
    m->normal[1] = nrm0;
  } else {
    v3 v1 = add(p1, axis1);
    real x2 = clampr((v - mb) / mc, -1, 1);
    v3 v2 = add(p2, scl(axis2, x2));
    collide_sphere_sphere(v1, r1, v2, r2, &m->dist[0], &m->pos[0], &nrm0);
    v3 n1;
    v1 = sub(p1, axis1);
    x2 = clampr((v + mb) / mc, -1, 1);
    v2 = add(p2, scl(axis2, x2));
    collide_sphere_sphere(v1, r1, v2, r2, &m->dist[1], &m->pos[1], &n1);
    m->normal[0] = nrm0;
    m->normal[1] = nrm0;
  }
}

/* collide_plane_capsule, :279-331 (contact frame unused by the writer) */
static void collide_plane_capsule(v3 n, v3 ppos, v3 cpos, v3 axis, real r, real hl, manifold* m) {
  v3 seg = scl(axis, hl);
  collide_plane_sphere(n, ppos, add(cpos, seg), r, &m->dist[0], &m->pos[0]);
  collide_plane_sphere(n, ppos, sub(cpos, seg), r, &m->dist[1], &m->pos[1]);
  m->normal[0] = n; m->normal[1] = n;
}

/* collide_plane_box, :366-439: keep up to 4 deepest corners within `margin` */
static void collide_plane_box(v3 n, v3 ppos, v3 bpos, const m33* rot, v3 size, real margin, manifold* m) {
  real center_dist = dot(sub(bpos, ppos), n);
  int nc = 0, worst = 0;
  for (int i = 0; i < 8; ++i) {
    v3 c = V((i & 1) ? size.x : -size.x, (i & 2) ? size.y : -size.y, (i & 4) ? size.z : -size.z);
    c = mv(rot, c);
    real cd = center_dist + dot(n, c);
    if (cd > margin) continue;
    v3 cpos = sub(add(c, bpos), scl(n, (real)0.5 * cd));
    if (nc < 4) {
      m->dist[nc] = cd; m->pos[nc] = cpos;
      if (nc == 0 || cd > m->dist[worst]) worst = nc;
      nc++;
    } else if (cd < m->dist[worst]) {
      m->dist[worst] = cd; m->pos[worst] = cpos;
      worst = 0;
      for (int k = 1; k < 4; ++k) if (m->dist[k] > m->dist[worst]) worst = k;
    }
  }
  for (int k = 0; k < 4; ++k) m->normal[k] = n;
}

/* collide_sphere_box, :1162-1214 */
static void collide_sphere_box(v3 sp, real sr, v3 bpos, const m33* rot, v3 size, real* dist, v3* pos, v3* n) {
  v3 center = mtv(rot, sub(sp, bpos));
  v3 cl = V(maxr(-size.x, minr(size.x, center.x)), maxr(-size.y, minr(size.y, center.y)),
            maxr(-size.z, minr(size.z, center.z)));
  v3 dv = sub(cl, center);
  real d = len(dv);
  v3 cdir = d > 0 ? scl(dv, (real)1 / d) : dv;
  v3 p, nl;
  if (d <= (real)1e-6) {
    real closest = 2 * (size.x + size.y + size.z);
    int k = 0;
    for (int i = 0; i < 6; ++i) {
      real fd = RA(((i % 2) ? (real)1 : (real)-1) * comp(size, i / 2) - comp(center, i / 2));
      if (closest > fd) { closest = fd; k = i; }
    }
    v3 nearest = V(0, 0, 0);
    setc(&nearest, k / 2, (k % 2) ? (real)-1 : (real)1);
    p = add(center, scl(nearest, (sr - closest) / 2));
    nl = nearest;
    *dist = -closest - sr;
  } else {
    v3 deepest = add(center, scl(cdir, sr));
    p = scl(add(cl, deepest), (real)0.5);
    nl = cdir;
    *dist = d - sr;
  }
  *n = mv(rot, nl);
  *pos = add(bpos, mv(rot, p));
}

/* ---- capsule-box (this project's restatement; reference uses GJK/MPR) ----
 * Signed distance of point p (box frame) to the box, and its outward gradient. */
static real box_sdf(v3 p, v3 h, v3* grad) {
  v3 q = V(RA(p.x) - h.x, RA(p.y) - h.y, RA(p.z) - h.z);
  v3 qo = V(maxr(q.x, 0), maxr(q.y, 0), maxr(q.z, 0));
  real out = len(qo);
  real in = minr(maxr(q.x, maxr(q.y, q.z)), 0);
  if (out > 0) {
    v3 g = V(p.x < 0 ? -qo.x : qo.x, p.y < 0 ? -qo.y : qo.y, p.z < 0 ? -qo.z : qo.z);
    *grad = scl(g, (real)1 / out);
  } else {
    int k = (q.x >= q.y && q.x >= q.z) ? 0 : (q.y >= q.z ? 1 : 2);
    v3 g = V(0, 0, 0);
    setc(&g, k, comp(p, k) < 0 ? (real)-1 : (real)1);
    *grad = g;
  }
  return out + in;
}

#define CAPBOX_ITERS 48
#define CAPBOX_PARALLEL 0.2

static void collide_capsule_box(v3 cpos, v3 axis, real r, real hl, v3 bpos, const m33* rot, v3 size, manifold* m) {
  v3 p0 = mtv(rot, sub(sub(cpos, scl(axis, hl)), bpos));
  v3 p1 = mtv(rot, sub(add(cpos, scl(axis, hl)), bpos));
  v3 d = sub(p1, p0);
  /* minimiser of the convex sdf(p0 + t d) over [0, 1]: bisection on the sign of the directional
   * derivative grad(sdf).d (accurate to rounding, where comparing values is only ~sqrt(eps)) */
  v3 g;
  real a = 0, b = 1;
  box_sdf(p0, size, &g);
  if (dot(g, d) >= 0) b = 0;
  else {
    box_sdf(add(p0, d), size, &g);
    if (dot(g, d) <= 0) a = 1;
    else
      for (int it = 0; it < CAPBOX_ITERS; ++it) {
        real mid = (real)0.5 * (a + b);
        box_sdf(add(p0, scl(d, mid)), size, &g);
        if (dot(g, d) < 0) a = mid; else b = mid;
      }
  }
  real ts = (real)0.5 * (a + b);
  v3 gs;
  real fs = box_sdf(add(p0, scl(d, ts)), size, &gs);
  real dl = len(d);
  /* Along a face the distance is linear in t over the part of the segment that lies above the
   * face, so its minimum is attained at an end of that part (and is not unique when the segment
   * is parallel).  Contacts: the two ends of the segment clipped to the face slab, plus the
   * golden-section minimum when it lies outside the clipped part (then it is unique). */
  int par = dl > 0 && RA(dot(d, gs)) < (real)CAPBOX_PARALLEL * dl;
  real tl[3];
  int nt = 0;
  if (par) {
    int fk = (RA(gs.x) >= RA(gs.y) && RA(gs.x) >= RA(gs.z)) ? 0 : (RA(gs.y) >= RA(gs.z) ? 1 : 2);
    real lo = 0, hi = 1;
    for (int j = 0; j < 3; ++j) {
      if (j == fk) continue;
      real pj = comp(p0, j), dj = comp(d, j), hj = comp(size, j);
      if (RA(dj) < (real)1e-12) {
        if (RA(pj) > hj) { lo = 1; hi = 0; }
      } else {
        real t1 = (-hj - pj) / dj, t2 = (hj - pj) / dj;
        lo = maxr(lo, minr(t1, t2));
        hi = minr(hi, maxr(t1, t2));
      }
    }
    if (lo <= hi) {
      tl[nt++] = lo;
      if (hi - lo > (real)1e-6) tl[nt++] = hi;
      v3 g_;
      real fe = minr(box_sdf(add(p0, scl(d, lo)), size, &g_), box_sdf(add(p0, scl(d, hi)), size, &g_));
      /* the minimum is strictly deeper than the clipped ends (off the face, or inside the box) */
      if (ts < lo - (real)1e-4 || ts > hi + (real)1e-4 || fs < fe - (real)1e-6) tl[nt++] = ts;
    } else {
      tl[nt++] = ts;
    }
  } else {
    tl[nt++] = ts;
  }
  for (int k = 0; k < nt; ++k) {
    v3 pk = add(p0, scl(d, tl[k]));
    v3 gk;
    real fk = box_sdf(pk, size, &gk);
    /* normal from capsule (A) to box (B) is -grad; midpoint between surfaces */
    v3 nl = scl(gk, -1);
    real dist = fk - r;
    v3 mid = add(pk, scl(nl, r + (real)0.5 * dist));
    m->dist[k] = dist;
    m->pos[k] = add(bpos, mv(rot, mid));
    m->normal[k] = mv(rot, nl);
  }
}

/* ---- box-box (this project's restatement; reference uses GJK/MPR + clipping) ---- */
#define BB_EDGE_BIAS 1.0e-4
static int clip_poly(const v3* in, int n, v3 pn, real off, v3* out) {
  /* keep the half-space dot(pn, x) <= off (Sutherland-Hodgman) */
  int m = 0;
  for (int i = 0; i < n; ++i) {
    v3 a = in[i], b = in[(i + 1) % n];
    real da = dot(pn, a) - off, db = dot(pn, b) - off;
    if (da <= 0) out[m++] = a;
    if ((da < 0 && db > 0) || (da > 0 && db < 0)) {
      real t = da / (da - db);
      out[m++] = add(a, scl(sub(b, a), t));
    }
  }
  return m;
}

static void collide_box_box(v3 pa, const m33* ra, v3 ha, v3 pb, const m33* rb, v3 hb, real gap, manifold* m) {
  v3 dv = sub(pb, pa);
  v3 A[3] = {col(ra, 0), col(ra, 1), col(ra, 2)}, B[3] = {col(rb, 0), col(rb, 1), col(rb, 2)};
  real hA[3] = {ha.x, ha.y, ha.z}, hB[3] = {hb.x, hb.y, hb.z};
  real best = -(real)OR_MAXVAL;
  int best_axis = -1;
  v3 best_n = V(0, 0, 1);
  real scale = maxr(maxr(ha.x, maxr(ha.y, ha.z)), maxr(hb.x, maxr(hb.y, hb.z)));
  for (int k = 0; k < 15; ++k) {
    v3 L;
    if (k < 3) L = A[k];
    else if (k < 6) L = B[k - 3];
    else {
      L = cross(A[(k - 6) / 3], B[(k - 6) % 3]);
      real l = len(L);
      if (l < (real)1e-6) continue;
      L = scl(L, (real)1 / l);
    }
    real rA = hA[0] * RA(dot(A[0], L)) + hA[1] * RA(dot(A[1], L)) + hA[2] * RA(dot(A[2], L));
    real rB = hB[0] * RA(dot(B[0], L)) + hB[1] * RA(dot(B[1], L)) + hB[2] * RA(dot(B[2], L));
    real dl = dot(dv, L);
    real s = RA(dl) - rA - rB;
    real bias = k < 6 ? 0 : (real)BB_EDGE_BIAS * scale;
    if (s > best + bias) { best = s; best_axis = k; best_n = dl < 0 ? scl(L, -1) : L; }
  }
  if (best > gap) return; /* separated beyond the contact gap on some axis */
  v3 n = best_n; /* from A to B */
  if (best_axis < 6) {
    int ref_is_a = best_axis < 3;
    const v3* R = ref_is_a ? A : B;
    const v3* I = ref_is_a ? B : A;
    const real* hR = ref_is_a ? hA : hB;
    const real* hI = ref_is_a ? hB : hA;
    v3 pR = ref_is_a ? pa : pb, pI = ref_is_a ? pb : pa;
    v3 nref = ref_is_a ? n : scl(n, -1); /* reference face normal, toward the incident box */
    int ax = ref_is_a ? best_axis : best_axis - 3;
    v3 fc = add(pR, scl(nref, hR[ax])); /* reference face centre */
    /* incident face: most anti-parallel to nref */
    int ik = 0; real mind = (real)OR_MAXVAL; real sgn = 1;
    for (int k = 0; k < 3; ++k) {
      real dd = dot(I[k], nref);
      if (dd < mind) { mind = dd; ik = k; sgn = 1; }
      if (-dd < mind) { mind = -dd; ik = k; sgn = -1; }
    }
    v3 ic = add(pI, scl(I[ik], sgn * hI[ik]));
    int u = (ik + 1) % 3, w = (ik + 2) % 3;
    v3 poly[8], tmp[8];
    poly[0] = add(ic, add(scl(I[u], hI[u]), scl(I[w], hI[w])));
    poly[1] = add(ic, add(scl(I[u], -hI[u]), scl(I[w], hI[w])));
    poly[2] = add(ic, add(scl(I[u], -hI[u]), scl(I[w], -hI[w])));
    poly[3] = add(ic, add(scl(I[u], hI[u]), scl(I[w], -hI[w])));
    int np = 4;
    int ru = (ax + 1) % 3, rw = (ax + 2) % 3;
    /* clip against the 4 side planes of the reference face */
    np = clip_poly(poly, np, R[ru], dot(R[ru], pR) + hR[ru], tmp);
    np = clip_poly(tmp, np, scl(R[ru], -1), -dot(R[ru], pR) + hR[ru], poly);
    np = clip_poly(poly, np, R[rw], dot(R[rw], pR) + hR[rw], tmp);
    np = clip_poly(tmp, np, scl(R[rw], -1), -dot(R[rw], pR) + hR[rw], poly);
    real dd[8]; v3 pp[8]; int kept = 0;
    for (int i = 0; i < np; ++i) {
      real di = dot(sub(poly[i], fc), nref);
      if (di > gap) continue;
      dd[kept] = di;
      pp[kept] = sub(poly[i], scl(nref, (real)0.5 * di)); /* midpoint to the reference face */
      kept++;
    }
    /* the whole clipped polygon is the manifold: it moves continuously with the poses, where a
     * reduction to 4 points would switch between near-equal-depth candidates */
    for (int i = 0; i < kept; ++i) {
      m->dist[i] = dd[i];
      m->pos[i] = pp[i];
      m->normal[i] = n;
    }
  } else {
    int ia = (best_axis - 6) / 3, ib = (best_axis - 6) % 3;
    /* supporting edges: A's edge toward +n, B's edge toward -n */
    v3 ca = pa, cb = pb;
    for (int k = 0; k < 3; ++k) {
      if (k != ia) ca = add(ca, scl(A[k], (dot(A[k], n) > 0 ? 1 : -1) * hA[k]));
      if (k != ib) cb = add(cb, scl(B[k], (dot(B[k], n) > 0 ? -1 : 1) * hB[k]));
    }
    /* closest points of the two edge lines, clamped to the edges */
    v3 da = A[ia], db = B[ib], r = sub(ca, cb);
    real b = dot(da, db), c = dot(da, r), f = dot(db, r);
    real den = 1 - b * b;
    real s = den > (real)1e-12 ? (b * f - c) / den : 0;
    s = clampr(s, -hA[ia], hA[ia]);
    real t = clampr(b * s + f, -hB[ib], hB[ib]);
    s = clampr(b * t - c, -hA[ia], hA[ia]);
    v3 qa = add(ca, scl(da, s)), qb = add(cb, scl(db, t));
    real di = dot(sub(qb, qa), n);
    m->dist[0] = di;
    m->pos[0] = scl(add(qa, qb), (real)0.5);
    m->normal[0] = n;
  }
}

/* ====================================================================
 * Shape poses, AABBs, broad phase, narrow phase
 * ==================================================================== */
typedef struct { v3 p; qt q; } xf;

static xf shape_world_xf(const or_model* M, const double* body_q, int b0, int s) {
  const double* sx = M->shape_xform + 7 * s;
  xf local = {V((real)sx[0], (real)sx[1], (real)sx[2]), {(real)sx[3], (real)sx[4], (real)sx[5], (real)sx[6]}};
  int b = M->shape_body[s];
  if (b < 0) return local;
  const double* bq = body_q + 7 * (b - b0);
  xf body = {V((real)bq[0], (real)bq[1], (real)bq[2]), {(real)bq[3], (real)bq[4], (real)bq[5], (real)bq[6]}};
  xf out;
  out.p = add(body.p, qrot(body.q, local.p));
  out.q = qmul(body.q, local.q);
  return out;
}

/* compute_shape_aabbs, collide.py:150-275 (closed forms of the support-map AABB) */
/* body_q holds the bodies starting at global index b0 */
void or_shape_aabb(const or_model* M, const double* body_q, int b0, int s, double* lo, double* hi) {
  xf X = shape_world_xf(M, body_q, b0, s);
  const double* sc = M->shape_scale + 3 * s;
  real eg = (real)(M->shape_margin[s] + M->shape_gap[s]);
  m33 R = qmat(X.q);
  v3 h;
  int t = M->shape_type[s];
  if (t == GEO_PLANE && sc[0] == 0.0 && sc[1] == 0.0) {
    real r = (real)M->shape_radius[s];
    h = V(r, r, r);
  } else if (t == GEO_SPHERE) {
    h = V((real)sc[0], (real)sc[0], (real)sc[0]);
  } else if (t == GEO_CAPSULE) {
    v3 ax = col(&R, 2);
    real hl = (real)sc[1], r = (real)sc[0];
    h = V(RA(ax.x) * hl + r, RA(ax.y) * hl + r, RA(ax.z) * hl + r);
  } else {
    v3 e = (t == GEO_PLANE) ? V((real)(0.5 * sc[0]), (real)(0.5 * sc[1]), 0) : V((real)sc[0], (real)sc[1], (real)sc[2]);
    h = V(RA(R.m[0]) * e.x + RA(R.m[1]) * e.y + RA(R.m[2]) * e.z, RA(R.m[3]) * e.x + RA(R.m[4]) * e.y + RA(R.m[5]) * e.z,
          RA(R.m[6]) * e.x + RA(R.m[7]) * e.y + RA(R.m[8]) * e.z);
  }
  lo[0] = X.p.x - h.x - eg; lo[1] = X.p.y - h.y - eg; lo[2] = X.p.z - h.z - eg;
  hi[0] = X.p.x + h.x + eg; hi[1] = X.p.y + h.y + eg; hi[2] = X.p.z + h.z + eg;
}

/* test_group_pair / test_world_and_group_pair, broad_phase_common.py:132-180 */
static int group_pair(int a, int b) {
  if (a == 0 || b == 0) return 0;
  if (a > 0) return a == b || b < 0;
  return a != b;
}
static int world_group_pair(int wa, int wb, int ga, int gb) {
  if (wa != -1 && wb != -1 && wa != wb) return 0;
  return group_pair(ga, gb);
}
static int excluded(const or_model* M, int a, int b) {
  int lo = 0, hi = M->filter_count - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int fa = M->filter_pairs[2 * mid], fb = M->filter_pairs[2 * mid + 1];
    if (fa == a && fb == b) return 1;
    if (a < fa || (a == fa && b < fb)) hi = mid - 1; else lo = mid + 1;
  }
  return 0;
}

/* Brute-force candidate pairs of one world segment (its shapes + the global shapes),
 * over the given AABBs (the n^2 oracle of test_broad_phase.py:90-148 with the pipeline's
 * zero extra cutoff, collide.py:1003-1016).  lo/hi are indexed by segment position
 * (world-local shapes first, then the global shapes).  Pairs are canonical (min, max) shape
 * indices, sorted lexicographically. */
static int cmp_pair(const void* a, const void* b) {
  const int32_t* x = (const int32_t*)a; const int32_t* y = (const int32_t*)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}
static int segment_shape(const or_model* M, int w, int i) {
  int s0 = M->world_shape_start[w], nl = M->world_shape_start[w + 1] - s0;
  return i < nl ? s0 + i : M->global_shapes[i - nl];
}
static int segment_size(const or_model* M, int w) {
  return M->world_shape_start[w + 1] - M->world_shape_start[w] + M->global_count;
}
int or_world_pairs(const or_model* M, int w, const double* lo, const double* hi, int32_t* pairs, int cap) {
  int n = segment_size(M, w);
  int cnt = 0;
  for (int i = 0; i < n; ++i) {
    int a = segment_shape(M, w, i);
    for (int j = i + 1; j < n; ++j) {
      int b = segment_shape(M, w, j);
      int sa = a < b ? a : b, sb = a < b ? b : a;
      int wa = M->shape_world[sa], wb = M->shape_world[sb];
      if (wa == -1 && wb == -1) continue; /* global-global pairs belong to the dedicated -1 segment */
      if (!world_group_pair(wa, wb, M->shape_group[sa], M->shape_group[sb])) continue;
      /* one body; the static shapes (body -1) count as one body too (builder.py:6204-6208) */
      if (M->shape_body[sa] == M->shape_body[sb]) continue;
      if (M->filter_count > 0 && excluded(M, sa, sb)) continue;
      const double *la = lo + 3 * i, *ua = hi + 3 * i, *lb = lo + 3 * j, *ub = hi + 3 * j;
      if (la[0] <= ub[0] && ua[0] >= lb[0] && la[1] <= ub[1] && ua[1] >= lb[1] && la[2] <= ub[2] && ua[2] >= lb[2]) {
        if (cnt < cap) { pairs[2 * cnt] = sa; pairs[2 * cnt + 1] = sb; }
        cnt++;
      }
    }
  }
  qsort(pairs, cnt < cap ? cnt : cap, 2 * sizeof(int32_t), cmp_pair);
  return cnt;
}

/* Test entry: pairs of world w from AABBs given per global shape index ([S,3] each). */
int or_world_pairs_global(const or_model* M, int w, const double* lo_g, const double* hi_g, int32_t* pairs, int cap) {
  int n = segment_size(M, w);
  double* lo = (double*)malloc(sizeof(double) * 3 * (n ? n : 1));
  double* hi = (double*)malloc(sizeof(double) * 3 * (n ? n : 1));
  for (int i = 0; i < n; ++i) {
    int s = segment_shape(M, w, i);
    memcpy(lo + 3 * i, lo_g + 3 * s, 3 * sizeof(double));
    memcpy(hi + 3 * i, hi_g + 3 * s, 3 * sizeof(double));
  }
  int cnt = or_world_pairs(M, w, lo, hi, pairs, cap);
  free(lo); free(hi);
  return cnt;
}

/* solver-side contact record (world frame at the collision state) */
typedef struct {
  int shape_a, shape_b, body_a, body_b, sub;
  real n[3], c[3];
  real dist, phi0, r_a, r_b, m_a, m_b, k, mu;
} or_contact;

/* Narrow phase of one canonical pair, narrow_phase.py:154-604 routing.
 * Returns the number of contacts written (gap-checked, contact_data.py:86-117). */
static int narrow_pair(const or_model* M, const double* body_q, int b0, int sa, int sb, or_contact* out) {
  int ta = M->shape_type[sa], tb = M->shape_type[sb];
  if (ta > tb) { int t = sa; sa = sb; sb = t; t = ta; ta = tb; tb = t; }
  xf Xa = shape_world_xf(M, body_q, b0, sa), Xb = shape_world_xf(M, body_q, b0, sb);
  const double *da = M->shape_scale + 3 * sa, *db = M->shape_scale + 3 * sb;
  v3 sca = V((real)da[0], (real)da[1], (real)da[2]), scb = V((real)db[0], (real)db[1], (real)db[2]);
  real gap_sum = (real)(M->shape_gap[sa] + M->shape_gap[sb]);
  real ma = (real)M->shape_margin[sa], mb = (real)M->shape_margin[sb];
  real ra = (ta == GEO_SPHERE || ta == GEO_CAPSULE) ? sca.x : 0;
  real rb = (tb == GEO_SPHERE || tb == GEO_CAPSULE) ? scb.x : 0;
  manifold m;
  mf_init(&m);
  m33 Ra = qmat(Xa.q), Rb = qmat(Xb.q);
  v3 za = col(&Ra, 2), zb = col(&Rb, 2);
  if (ta == GEO_PLANE && tb == GEO_SPHERE) {
    collide_plane_sphere(za, Xa.p, Xb.p, scb.x, &m.dist[0], &m.pos[0]);
    m.normal[0] = za;
  } else if (ta == GEO_PLANE && tb == GEO_BOX) {
    collide_plane_box(za, Xa.p, Xb.p, &Rb, scb, gap_sum, &m);
  } else if (ta == GEO_SPHERE && tb == GEO_SPHERE) {
    collide_sphere_sphere(Xa.p, sca.x, Xb.p, scb.x, &m.dist[0], &m.pos[0], &m.normal[0]);
  } else if (ta == GEO_PLANE && tb == GEO_CAPSULE) {
    collide_plane_capsule(za, Xa.p, Xb.p, zb, scb.x, scb.y, &m);
  } else if (ta == GEO_SPHERE && tb == GEO_CAPSULE) {
    collide_sphere_capsule(Xa.p, sca.x, Xb.p, zb, scb.x, scb.y, &m.dist[0], &m.pos[0], &m.normal[0]);
  } else if (ta == GEO_CAPSULE && tb == GEO_CAPSULE) {
    collide_capsule_capsule(Xa.p, za, sca.x, sca.y, Xb.p, zb, scb.x, scb.y, &m);
  } else if (ta == GEO_SPHERE && tb == GEO_BOX) {
    collide_sphere_box(Xa.p, sca.x, Xb.p, &Rb, scb, &m.dist[0], &m.pos[0], &m.normal[0]);
  } else if (ta == GEO_CAPSULE && tb == GEO_BOX) {
    collide_capsule_box(Xa.p, za, sca.x, sca.y, Xb.p, &Rb, scb, &m);
  } else if (ta == GEO_BOX && tb == GEO_BOX) {
    collide_box_box(Xa.p, &Ra, sca, Xb.p, &Rb, scb, gap_sum + ma + mb, &m);
  } else {
    return 0; /* plane-plane and unsupported types produce nothing */
  }
  int cnt = 0;
  for (int k = 0; k < OR_MAXMF; ++k) {
    if (!(m.dist[k] < (real)OR_MAXVAL)) continue;
    v3 n = nrm(m.normal[k]);
    /* contact_passes_gap_check: distance between the writer's core points minus radii+margins */
    v3 cpos = m.pos[k];
    v3 pa = sub(cpos, scl(n, (real)0.5 * m.dist[k] + ra));
    v3 pb = add(cpos, scl(n, (real)0.5 * m.dist[k] + rb));
    real d = dot(sub(pb, pa), n) - (ra + rb + ma + mb);
    if (!(d <= gap_sum)) continue;
    or_contact* c = out + cnt;
    c->shape_a = sa; c->shape_b = sb; c->sub = k;
    c->body_a = M->shape_body[sa]; c->body_b = M->shape_body[sb];
    c->n[0] = n.x; c->n[1] = n.y; c->n[2] = n.z;
    c->c[0] = cpos.x; c->c[1] = cpos.y; c->c[2] = cpos.z;
    c->dist = m.dist[k]; c->phi0 = d; c->r_a = ra; c->r_b = rb; c->m_a = ma; c->m_b = mb;
    c->k = (real)(0.5 * (M->shape_ke[sa] + M->shape_ke[sb]));
    c->mu = (real)(M->shape_mu[sa] > M->shape_mu[sb] ? M->shape_mu[sa] : M->shape_mu[sb]);
    cnt++;
  }
  return cnt;
}

/* Single-pair entry for the primitive KATs: shapes given directly (types, poses, scales).
 * out: [n*8] (n <= 8) = dist, pos(3), normal(3), 0 */
int or_collide_shapes(int ta, const double* xa, const double* sa_, int tb, const double* xb, const double* sb_,
                      double gap_sum, double* out) {
  xf Xa = {V((real)xa[0], (real)xa[1], (real)xa[2]), {(real)xa[3], (real)xa[4], (real)xa[5], (real)xa[6]}};
  xf Xb = {V((real)xb[0], (real)xb[1], (real)xb[2]), {(real)xb[3], (real)xb[4], (real)xb[5], (real)xb[6]}};
  v3 sca = V((real)sa_[0], (real)sa_[1], (real)sa_[2]), scb = V((real)sb_[0], (real)sb_[1], (real)sb_[2]);
  manifold m;
  mf_init(&m);
  m33 Ra = qmat(Xa.q), Rb = qmat(Xb.q);
  v3 za = col(&Ra, 2), zb = col(&Rb, 2);
  if (ta == GEO_PLANE && tb == GEO_SPHERE) { collide_plane_sphere(za, Xa.p, Xb.p, scb.x, &m.dist[0], &m.pos[0]); m.normal[0] = za; }
  else if (ta == GEO_PLANE && tb == GEO_BOX) collide_plane_box(za, Xa.p, Xb.p, &Rb, scb, (real)gap_sum, &m);
  else if (ta == GEO_SPHERE && tb == GEO_SPHERE) collide_sphere_sphere(Xa.p, sca.x, Xb.p, scb.x, &m.dist[0], &m.pos[0], &m.normal[0]);
  else if (ta == GEO_PLANE && tb == GEO_CAPSULE) collide_plane_capsule(za, Xa.p, Xb.p, zb, scb.x, scb.y, &m);
  else if (ta == GEO_SPHERE && tb == GEO_CAPSULE) collide_sphere_capsule(Xa.p, sca.x, Xb.p, zb, scb.x, scb.y, &m.dist[0], &m.pos[0], &m.normal[0]);
  else if (ta == GEO_CAPSULE && tb == GEO_CAPSULE) collide_capsule_capsule(Xa.p, za, sca.x, sca.y, Xb.p, zb, scb.x, scb.y, &m);
  else if (ta == GEO_SPHERE && tb == GEO_BOX) collide_sphere_box(Xa.p, sca.x, Xb.p, &Rb, scb, &m.dist[0], &m.pos[0], &m.normal[0]);
  else if (ta == GEO_CAPSULE && tb == GEO_BOX) collide_capsule_box(Xa.p, za, sca.x, sca.y, Xb.p, &Rb, scb, &m);
  else if (ta == GEO_BOX && tb == GEO_BOX) collide_box_box(Xa.p, &Ra, sca, Xb.p, &Rb, scb, (real)gap_sum, &m);
  int cnt = 0;
  for (int k = 0; k < OR_MAXMF; ++k) {
    if (!(m.dist[k] < (real)OR_MAXVAL)) continue;
    double* o = out + 8 * cnt;
    o[0] = m.dist[k]; o[1] = m.pos[k].x; o[2] = m.pos[k].y; o[3] = m.pos[k].z;
    o[4] = m.normal[k].x; o[5] = m.normal[k].y; o[6] = m.normal[k].z; o[7] = 0;
    cnt++;
  }
  return cnt;
}

/* Contact generation of one world at pose body_q (that world's bodies only):
 * AABBs -> brute-force pairs -> narrow phase.
 * Contacts are ordered by (shape_a, shape_b) of the canonical pair, then slot. */
/* Segment position of shape s in world w (local shapes in order, then the global shapes). */
static int segment_position(const or_model* M, int w, int s) {
  int s0 = M->world_shape_start[w], nl = M->world_shape_start[w + 1] - s0;
  if (s >= s0 && s < s0 + nl) return s - s0;
  for (int g = 0; g < M->global_count; ++g)
    if (M->global_shapes[g] == s) return nl + g;
  return 0;
}
/* Selection key of a contact when a world generates more than `cap` contacts (the device's
 * contact_key): phi0 as an order-preserving integer of its float32 value, then the pair's
 * segment positions (lower, higher) and the manifold slot.  Keys are unique within a world. */
static uint64_t contact_key(const or_model* M, int w, const or_contact* c) {
  float f = (float)c->phi0;
  uint32_t b;
  memcpy(&b, &f, 4);
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  int pa = segment_position(M, w, c->shape_a), pb = segment_position(M, w, c->shape_b);
  uint32_t lo = (uint32_t)(pa < pb ? pa : pb), hi = (uint32_t)(pa < pb ? pb : pa);
  return ((uint64_t)b << 32) | (lo << 16) | (hi << 8) | (uint32_t)c->sub;
}

/* Contacts of world w at its body poses, in pair order.  Returns the number generated; up to
 * `cap` are all kept, beyond it the set keeps the `cap` smallest keys (each further contact
 * replaces the largest key of the set when its own is smaller): the set the device's narrow phase
 * keeps, whose bulk selection per pass of pairs orders the slots differently. */
int or_collide_world(const or_model* M, int w, const double* body_q, or_contact* out, int cap, int* n_pairs) {
  int n = segment_size(M, w);
  double* lo = (double*)malloc(sizeof(double) * 3 * (n ? n : 1));
  double* hi = (double*)malloc(sizeof(double) * 3 * (n ? n : 1));
  int b0 = M->world_body_start[w];
  for (int i = 0; i < n; ++i) or_shape_aabb(M, body_q, b0, segment_shape(M, w, i), lo + 3 * i, hi + 3 * i);
  int pcap = n * (n - 1) / 2 + 1;
  int32_t* pairs = (int32_t*)malloc(sizeof(int32_t) * 2 * pcap);
  int np = or_world_pairs(M, w, lo, hi, pairs, pcap);
  if (n_pairs) *n_pairs = np;
  int cnt = 0;
  or_contact tmp[OR_MAXMF];
  for (int p = 0; p < np; ++p) {
    int k = narrow_pair(M, body_q, b0, pairs[2 * p], pairs[2 * p + 1], tmp);
    for (int i = 0; i < k; ++i) {
      if (cnt < cap) {
        out[cnt] = tmp[i];
      } else if (cap > 0) {
        uint64_t ck = contact_key(M, w, &tmp[i]), best = 0;
        int arg = 0;
        for (int c = 0; c < cap; ++c) {
          uint64_t k2 = contact_key(M, w, &out[c]);
          if (k2 > best) { best = k2; arg = c; }
        }
        if (ck < best) out[arg] = tmp[i];
      }
      cnt++;
    }
  }
  free(pairs); free(lo); free(hi);
  return cnt;
}

/* Flat export of one world's contacts for tests: per contact 14 doubles
 * (shape_a, shape_b, sub, n[3], c[3], dist, phi0, k, mu, r_a) */
int or_collide_world_flat(const or_model* M, int w, const double* body_q, double* out, int cap, int* n_pairs) {
  or_contact* cs = (or_contact*)malloc(sizeof(or_contact) * (cap > 0 ? cap : 1));
  int cnt = or_collide_world(M, w, body_q + 7 * M->world_body_start[w], cs, cap, n_pairs);
  int m = cnt < cap ? cnt : cap;
  for (int i = 0; i < m; ++i) {
    double* o = out + 14 * i;
    o[0] = cs[i].shape_a; o[1] = cs[i].shape_b; o[2] = cs[i].sub;
    o[3] = cs[i].n[0]; o[4] = cs[i].n[1]; o[5] = cs[i].n[2];
    o[6] = cs[i].c[0]; o[7] = cs[i].c[1]; o[8] = cs[i].c[2];
    o[9] = cs[i].dist; o[10] = cs[i].phi0; o[11] = cs[i].k; o[12] = cs[i].mu; o[13] = cs[i].r_a;
  }
  free(cs);
  return cnt;
}

/* ====================================================================
 * Convex SAP solve of one world (dense, exact Newton).
 * ==================================================================== */
typedef struct {
  int nb, nc, N;
  real* m;     /* [nb] */
  real* Iw;    /* [nb*9] world inertia */
  real* Iwi;   /* [nb*9] its inverse */
  real* xc;    /* [nb*3] COM position */
  acc* vstar; /* [N] */
  acc* Ad;    /* [N] diag(A) */
  /* contacts */
  int* cb;     /* [nc*2] local body indices (-1 static) */
  real* J;     /* [nc*2*18] 3x6 blocks per body (rows t1, t2, n) */
  real* Rt; real* Rn; real* vhat; real* mu;
  /* scratch */
  real* H; acc* g; acc* dv; acc* v; acc* gam; real* G; acc* tmp; real* rtmp;
  /* generalized-coordinate problems (articulated worlds): dense A [N,N] and dense contact rows
     [nc, 3, N]; NULL for free bodies (block-diagonal A, per-body 3x6 blocks) */
  real* Adense;
  real* Jd;
  /* contact law 1 (MuJoCo): per contact the regulariser D = 1/R of every row (0: inactive) and
     the contact-frame velocity the rows are driven towards, what [3 nc] (t1, t2, n) */
  int law;
  real* D; real* what;
} sap_ctx;

/* P_F projection in the R norm and its Jacobian-derived Hessian G = dP/dy R^-1 */
static acc project(acc mu, acc Rt, acc Rn, const acc* y, acc* gam, real* G) {
  acc yr = sqrt(y[0] * y[0] + y[1] * y[1]), yn = y[2];
  acc muh = mu * Rt / Rn;
  for (int i = 0; i < 9; ++i) G[i] = 0;
  if (yr <= mu * yn) {
    gam[0] = y[0]; gam[1] = y[1]; gam[2] = y[2];
    G[0] = (real)(1 / Rt); G[4] = (real)(1 / Rt); G[8] = (real)(1 / Rn);
  } else if (yn <= -muh * yr) {
    gam[0] = gam[1] = gam[2] = 0;
  } else {
    acc t0 = y[0] / yr, t1 = y[1] / yr;
    /* OR_PLANT (tests only, never in the default build): plausible mistakes in the sliding branch
       that the friction pins must catch -- 1 drops mu mu_hat from the denominator, 2 drops the
       mu_hat |y_t| term of the normal impulse, 3 scales the tangential impulse by mu^2 */
#if defined(OR_PLANT) && OR_PLANT == 1
    acc den = 1;
#else
    acc den = 1 + mu * muh;
#endif
#if defined(OR_PLANT) && OR_PLANT == 2
    acc gn = yn / den;
#else
    acc gn = (yn + muh * yr) / den;
#endif
#if defined(OR_PLANT) && OR_PLANT == 3
    gam[0] = mu * mu * gn * t0; gam[1] = mu * mu * gn * t1; gam[2] = gn;
#else
    gam[0] = mu * gn * t0; gam[1] = mu * gn * t1; gam[2] = gn;
#endif
    acc a = mu * muh / den, b = mu * gn / yr;
    acc P[9];
    P[0] = a * t0 * t0 + b * (1 - t0 * t0); P[1] = a * t0 * t1 - b * t0 * t1; P[2] = mu / den * t0;
    P[3] = a * t1 * t0 - b * t1 * t0;       P[4] = a * t1 * t1 + b * (1 - t1 * t1); P[5] = mu / den * t1;
    P[6] = muh / den * t0;                  P[7] = muh / den * t1;               P[8] = 1 / den;
    for (int r = 0; r < 3; ++r) {
      G[3 * r + 0] = (real)(P[3 * r + 0] / Rt); G[3 * r + 1] = (real)(P[3 * r + 1] / Rt);
      G[3 * r + 2] = (real)(P[3 * r + 2] / Rn);
    }
  }
  return 0.5 * (Rt * (gam[0] * gam[0] + gam[1] * gam[1]) + Rn * gam[2] * gam[2]);
}

/* ---------------------------------------------------------------- MuJoCo soft constraints
 * The soft-constraint model of MuJoCo's convex solver (the inner step of SolverMuJoCoAdaptive,
 * mujoco_warp.step, solver_mujoco_adaptive.py:585-605; absent dependency mujoco-warp 3.8.1),
 * restated from MuJoCo's published computation model: each constraint row i has a reference
 * acceleration aref_i = -B (J_i v) - K imp (pos - margin), a regulariser R_i = (1 - imp)/imp
 * invweight_i and, for contacts, the one-sided cost 1/2 R_i^-1 min(J_i a - aref_i, 0)^2; the
 * primal problem min_a 1/2 |a - a0|_M^2 + sum_i cost_i is solved and v' = v + h a (Euler).
 * In velocity form (a = (v' - v)/h, scaled by h^2) a row costs 1/2 D min(J_i v' - vhat_i, 0)^2
 * with D = 1/R and vhat_i = J_i v + h aref_i.  A pyramidal contact of condim 3 has four rows
 * J_n +- mu J_tk, all with the invweight (1 + mu^2) w 2 mu^2 / impratio (w: the two bodies'
 * translational inverse weights); condim 1 has the normal row alone with invweight w. */
#define OR_MJ_MINVAL 1.0e-15
#define OR_MJ_MINIMP 0.0001
#define OR_MJ_MAXIMP 0.9999
#define OR_MJ_MINMU 1.0e-5

/* impedance d(pos) of solimp = (dmin, dmax, width, midpoint, power): the two-piece power curve
   from dmin at |pos| = 0 to dmax at |pos| >= width (MuJoCo's solimp definition) */
double or_mj_impedance(const double* solimp, double pos) {
  double dmin = clampr(solimp[0], OR_MJ_MINIMP, OR_MJ_MAXIMP), dmax = clampr(solimp[1], OR_MJ_MINIMP, OR_MJ_MAXIMP);
  double width = maxr(OR_MJ_MINVAL, solimp[2]), mid = clampr(solimp[3], OR_MJ_MINIMP, OR_MJ_MAXIMP);
  double power = maxr(1.0, solimp[4]);
  double x = fabs(pos) / width;
  if (x >= 1.0) return dmax;
  double y;
  if (power == 1.0) y = x;
  else if (x < mid) y = pow(x, power) / pow(mid, power - 1.0);
  else y = 1.0 - pow(1.0 - x, power) / pow(1.0 - mid, power - 1.0);
  return clampr(dmin + y * (dmax - dmin), dmin, dmax);
}

/* Contact parameters of a shape pair (contact_params, solvers/mujoco/kernels.py:127-190): the
   higher priority shape decides; equal priorities mix solref/solimp by solmix (0.5 when both are
   ~0, one-sided when one is), take the larger condim and the larger friction; a non-positive
   solref time constant on either side selects the elementwise min.  Friction floored at MJ_MINMU. */
void or_mj_contact_params(const or_model* M, int a, int b, double* solref, double* solimp, double* mu, int* condim) {
  int pa = M->mj_priority[a], pb = M->mj_priority[b];
  double mix;
  if (pa > pb) { mix = 1.0; *condim = M->mj_condim[a]; *mu = M->shape_mu[a]; }
  else if (pb > pa) { mix = 0.0; *condim = M->mj_condim[b]; *mu = M->shape_mu[b]; }
  else {
    double s1 = M->mj_solmix[a], s2 = M->mj_solmix[b];
    const double tiny = 2.220446049250313e-16;
    mix = s1 + s2 != 0.0 ? s1 / (s1 + s2) : 0.0;
    if (s1 < tiny && s2 < tiny) mix = 0.5;
    else if (s1 < tiny) mix = 0.0;
    else if (s2 < tiny) mix = 1.0;
    *condim = M->mj_condim[a] > M->mj_condim[b] ? M->mj_condim[a] : M->mj_condim[b];
    *mu = maxr(M->shape_mu[a], M->shape_mu[b]);
  }
  *mu = maxr(OR_MJ_MINMU, *mu);
  const double *ra = M->mj_solref + 2 * a, *rb = M->mj_solref + 2 * b;
  for (int k = 0; k < 2; ++k)
    solref[k] = (ra[0] > 0 && rb[0] > 0) ? mix * ra[k] + (1 - mix) * rb[k] : minr(ra[k], rb[k]);
  for (int k = 0; k < 5; ++k) solimp[k] = mix * M->mj_solimp[5 * a + k] + (1 - mix) * M->mj_solimp[5 * b + k];
}

/* Cost of one contact's rows at y = w - what (contact frame t1, t2, n), its impulse gamma
   (minus the gradient in w) and Hessian G (3x3).  mu == 0: the normal row alone (condim 1). */
static acc mj_pyramid(acc mu, acc D, const acc* y, acc* gam, real* G) {
  acc cost = 0;
  gam[0] = gam[1] = gam[2] = 0;
  for (int i = 0; i < 9; ++i) G[i] = 0;
  if (D <= 0) return 0;
  if (mu == 0) {
    if (y[2] < 0) { cost = 0.5 * D * y[2] * y[2]; gam[2] = -D * y[2]; G[8] = (real)D; }
    return cost;
  }
  acc Gd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 2; ++k) {
    for (int s = -1; s <= 1; s += 2) {
#if defined(OR_PLANT) && OR_PLANT == 4
      acc x = y[2] + s * y[k];
#else
      acc x = y[2] + s * mu * y[k];
#endif
      if (!(x < 0)) continue;
      cost += 0.5 * D * x * x;
      gam[2] -= D * x;
      gam[k] -= D * x * s * mu;
      Gd[8] += D;
      Gd[4 * k] += D * mu * mu;
      Gd[3 * k + 2] += D * s * mu;
      Gd[6 + k] += D * s * mu;
    }
  }
  for (int i = 0; i < 9; ++i) G[i] = (real)Gd[i];
  return cost;
}

/* cost, impulse and Hessian of contact c at contact velocity vc under the context's law */
static acc law_eval(const sap_ctx* S, int c, const acc* vc, acc* gam, real* G) {
  acc y[3];
  if (S->law == 1) {
    for (int k = 0; k < 3; ++k) y[k] = vc[k] - S->what[3 * c + k];
    return mj_pyramid(S->mu[c], S->D[c], y, gam, G);
  }
  y[0] = -vc[0] / S->Rt[c]; y[1] = -vc[1] / S->Rt[c]; y[2] = -(vc[2] - S->vhat[c]) / S->Rn[c];
  return project(S->mu[c], S->Rt[c], S->Rn[c], y, gam, G);
}

static void contact_velocity(const sap_ctx* S, int c, const acc* v, acc* vc) {
  vc[0] = vc[1] = vc[2] = 0;
  if (S->Jd) {
    for (int r = 0; r < 3; ++r) {
      const real* row = S->Jd + ((size_t)3 * c + r) * S->N;
      acc sum = 0;
      for (int k = 0; k < S->N; ++k) sum += row[k] * v[k];
      vc[r] = sum;
    }
    return;
  }
  for (int s = 0; s < 2; ++s) {
    int b = S->cb[2 * c + s];
    if (b < 0) continue;
    const real* Jb = S->J + (2 * c + s) * 18;
    const acc* vb = v + 6 * b;
    for (int r = 0; r < 3; ++r) {
      acc sum = 0;
      for (int k = 0; k < 6; ++k) sum += Jb[6 * r + k] * vb[k];
      vc[r] += sum;
    }
  }
}

/* momentum A x for block-diagonal A = diag(m I, Iw) */
static void apply_A(const sap_ctx* S, const acc* x, acc* y) {
  if (S->Adense) {
    for (int i = 0; i < S->N; ++i) {
      acc sum = 0;
      for (int k = 0; k < S->N; ++k) sum += S->Adense[(size_t)i * S->N + k] * x[k];
      y[i] = sum;
    }
    return;
  }
  for (int b = 0; b < S->nb; ++b) {
    const acc* xb = x + 6 * b; acc* yb = y + 6 * b; const real* I = S->Iw + 9 * b;
    yb[0] = S->m[b] * xb[0]; yb[1] = S->m[b] * xb[1]; yb[2] = S->m[b] * xb[2];
    for (int r = 0; r < 3; ++r) yb[3 + r] = I[3 * r] * xb[3] + I[3 * r + 1] * xb[4] + I[3 * r + 2] * xb[5];
  }
}

/* cost, gradient (g = A(v-v*) - J^T gamma) and, if G != NULL, per-contact Hessian blocks */
static acc eval_cost(sap_ctx* S, const acc* v, acc* g, acc* gam, real* G, acc* pnorm, acc* jnorm) {
  int N = S->N;
  acc* d = S->tmp;
  for (int i = 0; i < N; ++i) d[i] = v[i] - S->vstar[i];
  apply_A(S, d, g);
  acc cost = 0;
  for (int i = 0; i < N; ++i) cost += 0.5 * d[i] * g[i];
  acc* jc = S->tmp + N;
  for (int i = 0; i < N; ++i) jc[i] = 0;
  for (int c = 0; c < S->nc; ++c) {
    acc vc[3], gl[3];
    real Gl[9];
    contact_velocity(S, c, v, vc);
    cost += law_eval(S, c, vc, gl, Gl);
    if (gam) memcpy(gam + 3 * c, gl, sizeof(gl));
    if (G) memcpy(G + 9 * c, Gl, sizeof(Gl));
    if (S->Jd) {
      for (int r = 0; r < 3; ++r) {
        const real* row = S->Jd + ((size_t)3 * c + r) * N;
        for (int k = 0; k < N; ++k) jc[k] += row[k] * gl[r];
      }
      continue;
    }
    for (int s = 0; s < 2; ++s) {
      int b = S->cb[2 * c + s];
      if (b < 0) continue;
      const real* Jb = S->J + (2 * c + s) * 18;
      for (int k = 0; k < 6; ++k) jc[6 * b + k] += Jb[k] * gl[0] + Jb[6 + k] * gl[1] + Jb[12 + k] * gl[2];
    }
  }
  for (int i = 0; i < N; ++i) g[i] -= jc[i];
  if (pnorm) {
    acc* p = S->tmp + 2 * N;
    apply_A(S, v, p);
    acc pn = 0, jn = 0;
    for (int i = 0; i < N; ++i) { acc s = 1 / sqrt(S->Ad[i]); pn += (s * p[i]) * (s * p[i]); jn += (s * jc[i]) * (s * jc[i]); }
    *pnorm = sqrt(pn); *jnorm = sqrt(jn);
  }
  return cost;
}

/* l(alpha) along dv: value, first and second derivative */
static void ls_eval(sap_ctx* S, const acc* v, const acc* dv, acc alpha, acc* l, acc* dl, acc* d2l) {
  int N = S->N;
  acc* d = S->tmp;
  acc* Adv = S->tmp + N;
  for (int i = 0; i < N; ++i) d[i] = v[i] + alpha * dv[i] - S->vstar[i];
  apply_A(S, dv, Adv);
  acc mom = 0, dm = 0, ddm = 0;
  acc* Ad = S->tmp + 2 * N;
  apply_A(S, d, Ad);
  for (int i = 0; i < N; ++i) { mom += 0.5 * d[i] * Ad[i]; dm += dv[i] * Ad[i]; ddm += dv[i] * Adv[i]; }
  acc* va = S->tmp + 3 * N;
  for (int i = 0; i < N; ++i) va[i] = v[i] + alpha * dv[i];
  for (int c = 0; c < S->nc; ++c) {
    acc vc[3], dvc[3], gl[3];
    real Gl[9];
    contact_velocity(S, c, va, vc);
    contact_velocity(S, c, dv, dvc);
    mom += law_eval(S, c, vc, gl, Gl);
    dm -= gl[0] * dvc[0] + gl[1] * dvc[1] + gl[2] * dvc[2];
    for (int r = 0; r < 3; ++r) ddm += dvc[r] * (Gl[3 * r] * dvc[0] + Gl[3 * r + 1] * dvc[1] + Gl[3 * r + 2] * dvc[2]);
  }
  *l = mom; *dl = dm; *d2l = ddm;
}

static int cholesky(real* A, int N) {
  for (int j = 0; j < N; ++j) {
    real s = A[j * N + j];
    for (int k = 0; k < j; ++k) s -= A[j * N + k] * A[j * N + k];
    if (!(s > 0)) return 0;
    real d = RS(s);
    A[j * N + j] = d;
    for (int i = j + 1; i < N; ++i) {
      real t = A[i * N + j];
      for (int k = 0; k < j; ++k) t -= A[i * N + k] * A[j * N + k];
      A[i * N + j] = t / d;
    }
  }
  return 1;
}
static void chol_solve(const real* L, int N, real* x) {
  for (int i = 0; i < N; ++i) { real s = x[i]; for (int k = 0; k < i; ++k) s -= L[i * N + k] * x[k]; x[i] = s / L[i * N + i]; }
  for (int i = N - 1; i >= 0; --i) { real s = x[i]; for (int k = i + 1; k < N; ++k) s -= L[k * N + i] * x[k]; x[i] = s / L[i * N + i]; }
}

typedef struct { int iterations, factorizations, converged; acc residual; } solve_stats;

/* Decrease l(v + alpha dv) - l(v) of the line function, summed term by term: the momentum part in
   closed form (alpha dv^T A (v - v*) + alpha^2 / 2 dv^T A dv), the contact costs as per-contact
   differences, so that it is not lost in the rounding of l itself. */
static acc ls_delta(sap_ctx* S, const acc* v, const acc* dv, acc alpha) {
  int N = S->N;
  acc* d = S->tmp;
  acc* Ad = S->tmp + N;
  acc* Adv = S->tmp + 2 * N;
  acc* va = S->tmp + 3 * N;
  for (int i = 0; i < N; ++i) d[i] = v[i] - S->vstar[i];
  apply_A(S, d, Ad);
  apply_A(S, dv, Adv);
  acc b0 = 0, aa = 0;
  for (int i = 0; i < N; ++i) { b0 += dv[i] * Ad[i]; aa += dv[i] * Adv[i]; }
  acc delta = alpha * b0 + 0.5 * alpha * alpha * aa;
  for (int i = 0; i < N; ++i) va[i] = v[i] + alpha * dv[i];
  for (int c = 0; c < S->nc; ++c) {
    acc vc0[3], vca[3], gl[3];
    real Gl[9];
    contact_velocity(S, c, v, vc0);
    contact_velocity(S, c, va, vca);
    acc ca = law_eval(S, c, vca, gl, Gl);
    acc c0 = law_eval(S, c, vc0, gl, Gl);
    delta += ca - c0;
  }
  return delta;
}

/* Backtracking line search with Armijo's sufficient decrease and geometric decay along S->dv from
   v: the "armijo_decay" variant of the reference's SolverSAPAdaptive (solver_sap_adaptive.py:459;
   its solver sap_warp is absent), restated from Drake's SAP backtracking search (alpha_max 1.5,
   rho 0.8, c = 1e-4, 40 evaluations).  Steps alpha_max rho^k are tried in turn; the search stops at
   the first step whose cost rises above the previous one's once the previous one met
   l(a) - l(0) < c a l'(0), else ends with the last step if it meets that condition.  Returns the
   step, or -1 when no evaluated step met it (near the optimum the decrease can fall below the
   rounding of the contact costs; the caller then takes the exact search's step). */
static acc armijo_decay(sap_ctx* S, const acc* v, acc dl0) {
  const acc alpha_max = 1.5, rho = 0.8, c = 1.0e-4;
  acc alpha = alpha_max;
  acc d = ls_delta(S, v, S->dv, alpha);
  int armijo = d < c * alpha * dl0;
  acc alpha_prev = alpha, d_prev = d;
  for (int k = 1; k < 40; ++k) {
    alpha *= rho;
    d = ls_delta(S, v, S->dv, alpha);
    if (d > d_prev && armijo) return alpha_prev;
    armijo = d < c * alpha * dl0;
    alpha_prev = alpha;
    d_prev = d;
  }
  return armijo ? alpha : -1.0;
}

static int sap_newton(sap_ctx* S, const or_config* C, acc* v, solve_stats* st) {
  int N = S->N;
  acc rel = C->opt_rel_tol, abs_tol = C->opt_abs_tol;
  st->iterations = 0; st->factorizations = 0; st->converged = 0;
  for (int it = 0; it <= C->max_iterations; ++it) {
    acc pn, jn;
    eval_cost(S, v, S->g, S->gam, S->G, &pn, &jn);
    acc gn = 0;
    for (int i = 0; i < N; ++i) { acc s = S->g[i] / sqrt(S->Ad[i]); gn += s * s; }
    gn = sqrt(gn);
    st->residual = gn;
    if (!(gn == gn)) return 0;
    if (gn <= abs_tol + rel * maxr(pn, jn)) { st->converged = 1; return 1; }
    if (it == C->max_iterations) break;
    /* H = A + sum J^T G J (dense) */
    memset(S->H, 0, sizeof(real) * N * N);
    if (S->Adense) {
      memcpy(S->H, S->Adense, sizeof(real) * N * N);
      for (int c = 0; c < S->nc; ++c) {
        const real* Gc = S->G + 9 * c;
        const real* Jc = S->Jd + (size_t)3 * c * N;
        for (int i = 0; i < N; ++i) {
          real gj[3];
          for (int r = 0; r < 3; ++r) gj[r] = Gc[3 * r] * Jc[i] + Gc[3 * r + 1] * Jc[N + i] + Gc[3 * r + 2] * Jc[2 * N + i];
          for (int k = 0; k < N; ++k) S->H[(size_t)k * N + i] += Jc[k] * gj[0] + Jc[N + k] * gj[1] + Jc[2 * N + k] * gj[2];
        }
      }
    }
    for (int b = 0; b < S->nb && !S->Adense; ++b) {
      int o = 6 * b;
      S->H[o * N + o] = S->m[b]; S->H[(o + 1) * N + o + 1] = S->m[b]; S->H[(o + 2) * N + o + 2] = S->m[b];
      for (int r = 0; r < 3; ++r) for (int q = 0; q < 3; ++q) S->H[(o + 3 + r) * N + o + 3 + q] = S->Iw[9 * b + 3 * r + q];
    }
    for (int c = 0; c < S->nc && !S->Adense; ++c) {
      const real* Gc = S->G + 9 * c;
      for (int s1 = 0; s1 < 2; ++s1) {
        int b1 = S->cb[2 * c + s1];
        if (b1 < 0) continue;
        const real* J1 = S->J + (2 * c + s1) * 18;
        real GJ[18]; /* G * J1 (3x6) */
        for (int r = 0; r < 3; ++r) for (int k = 0; k < 6; ++k) GJ[6 * r + k] = Gc[3 * r] * J1[k] + Gc[3 * r + 1] * J1[6 + k] + Gc[3 * r + 2] * J1[12 + k];
        for (int s2 = 0; s2 < 2; ++s2) {
          int b2 = S->cb[2 * c + s2];
          if (b2 < 0) continue;
          const real* J2 = S->J + (2 * c + s2) * 18;
          for (int i = 0; i < 6; ++i) for (int k = 0; k < 6; ++k)
            S->H[(6 * b2 + i) * N + 6 * b1 + k] += J2[i] * GJ[k] + J2[6 + i] * GJ[6 + k] + J2[12 + i] * GJ[12 + k];
        }
      }
    }
    if (!cholesky(S->H, N)) return 0;
    st->factorizations++;
    /* the direction from the (real) factor; the iterate, gradient and line search stay in acc */
    for (int i = 0; i < N; ++i) S->rtmp[i] = (real)(-S->g[i]);
    chol_solve(S->H, N, S->rtmp);
    for (int i = 0; i < N; ++i) S->dv[i] = S->rtmp[i];
    /* line search: exact on [0, 1] (root of l'(alpha), cubic-seeded safeguarded Newton) or
       backtracking with Armijo's condition ("armijo_decay") */
    acc l0, d0, dd0, l1, d1, dd1;
    ls_eval(S, v, S->dv, 0, &l0, &d0, &dd0);
    if (!(d0 < 0)) { st->iterations = it + 1; break; }
    acc alpha = -1;
    if (C->line_search == 1) alpha = armijo_decay(S, v, d0);
    ls_eval(S, v, S->dv, 1, &l1, &d1, &dd1);
    if (alpha > 0) {
      /* backtracking step taken */
    } else if (d1 <= 0) alpha = 1;
    else {
      acc lo = 0, hi = 1;
      double ac;
      alpha = or_cubic_min(l0, d0, l1, d1, 1.0, &ac) ? ac : fmin(fmax(-d0 / dd0, 0.0), 1.0);
      if (!(alpha > lo && alpha < hi)) alpha = 0.5;
      for (int k = 0; k < 100; ++k) {
        acc l, d, dd;
        ls_eval(S, v, S->dv, alpha, &l, &d, &dd);
        if (fabs(d) <= C->ls_rel_tol * fabs(d0)) break;
        if (d < 0) lo = alpha; else hi = alpha;
        acc an = dd > 0 ? alpha - d / dd : -1;
        if (!(an > lo && an < hi)) an = 0.5 * (lo + hi);
        if (hi - lo <= 1e-14 * hi) { alpha = an; break; }
        alpha = an;
      }
    }
    for (int i = 0; i < N; ++i) v[i] += alpha * S->dv[i];
    st->iterations = it + 1;
  }
  return st->converged;
}

/* ---- per-world state views and integration ---- */
typedef struct {
  const or_model* M;
  int w, b0, nb;
} world_view;

/* Build the SAP problem of one world for step h from pose q, velocity qd, wrench f. */
static void sap_setup(sap_ctx* S, const world_view* W, const or_config* C, const double* q, const double* qd,
                      const double* f, real h, const or_contact* cs, int nc) {
  const or_model* M = W->M;
  int nb = W->nb;
  S->nb = nb; S->nc = nc; S->N = 6 * nb;
  for (int b = 0; b < nb; ++b) {
    int gb = W->b0 + b;
    const double* bq = q + 7 * b;
    qt r = {(real)bq[3], (real)bq[4], (real)bq[5], (real)bq[6]};
    m33 R = qmat(r);
    const double* Ib = M->body_inertia + 9 * gb;
    const double* Ibi = M->body_inv_inertia + 9 * gb;
    real m = (real)M->body_mass[gb];
    S->m[b] = m;
    /* Iw = R Ib R^T and its inverse R Ib^-1 R^T */
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
      real a = 0, ai = 0;
      for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) {
        a += R.m[3 * i + k] * (real)Ib[3 * k + l] * R.m[3 * j + l];
        ai += R.m[3 * i + k] * (real)Ibi[3 * k + l] * R.m[3 * j + l];
      }
      S->Iw[9 * b + 3 * i + j] = a; S->Iwi[9 * b + 3 * i + j] = ai;
    }
    v3 com = V((real)M->body_com[3 * gb], (real)M->body_com[3 * gb + 1], (real)M->body_com[3 * gb + 2]);
    v3 xc = add(V((real)bq[0], (real)bq[1], (real)bq[2]), mv(&R, com));
    S->xc[3 * b] = xc.x; S->xc[3 * b + 1] = xc.y; S->xc[3 * b + 2] = xc.z;
    /* free motion: v* = v0 + h (g + f/m),  w* = w0 + h Iw^-1 (tau - w0 x Iw w0) */
    const double* vb = qd + 6 * b;
    const double* fb = f + 6 * b;
    real im = (real)M->body_inv_mass[gb];
    for (int k = 0; k < 3; ++k) S->vstar[6 * b + k] = (real)vb[k] + h * ((real)fb[k] * im + (real)M->gravity[k] * (im > 0 ? 1 : 0));
    v3 w0 = V((real)vb[3], (real)vb[4], (real)vb[5]);
    v3 Iw0 = V(S->Iw[9 * b] * w0.x + S->Iw[9 * b + 1] * w0.y + S->Iw[9 * b + 2] * w0.z,
               S->Iw[9 * b + 3] * w0.x + S->Iw[9 * b + 4] * w0.y + S->Iw[9 * b + 5] * w0.z,
               S->Iw[9 * b + 6] * w0.x + S->Iw[9 * b + 7] * w0.y + S->Iw[9 * b + 8] * w0.z);
    v3 tau = sub(V((real)fb[3], (real)fb[4], (real)fb[5]), cross(w0, Iw0));
    const real* Ii = S->Iwi + 9 * b;
    S->vstar[6 * b + 3] = w0.x + h * (Ii[0] * tau.x + Ii[1] * tau.y + Ii[2] * tau.z);
    S->vstar[6 * b + 4] = w0.y + h * (Ii[3] * tau.x + Ii[4] * tau.y + Ii[5] * tau.z);
    S->vstar[6 * b + 5] = w0.z + h * (Ii[6] * tau.x + Ii[7] * tau.y + Ii[8] * tau.z);
    S->Ad[6 * b] = S->Ad[6 * b + 1] = S->Ad[6 * b + 2] = m;
    S->Ad[6 * b + 3] = S->Iw[9 * b]; S->Ad[6 * b + 4] = S->Iw[9 * b + 4]; S->Ad[6 * b + 5] = S->Iw[9 * b + 8];
  }
  real tau_d = (real)C->tau_d;
  S->law = C->contact_law;
  for (int c = 0; c < nc; ++c) {
    const or_contact* k = cs + c;
    v3 n = V(k->n[0], k->n[1], k->n[2]);
    v3 t1;
    if (C->contact_law == 1) {
      /* MuJoCo's contact frame (orthogonals, solvers/mujoco/kernels.py:47-58): t1 from the y axis,
         or the z axis when |n_y| >= 1/2, made orthogonal to n */
      v3 b = (-0.5f < n.y && n.y < 0.5f) ? V(0, 1, 0) : V(0, 0, 1);
      t1 = nrm(sub(b, scl(n, dot(n, b))));
    } else {
      /* tangent basis: t1 orthogonal to n using the axis least aligned with n */
      v3 e = RA(n.x) < RA(n.y) ? (RA(n.x) < RA(n.z) ? V(1, 0, 0) : V(0, 0, 1)) : (RA(n.y) < RA(n.z) ? V(0, 1, 0) : V(0, 0, 1));
      t1 = nrm(cross(n, e));
    }
    v3 t2 = cross(n, t1);
    v3 F[3] = {t1, t2, n};
    v3 p = V(k->c[0], k->c[1], k->c[2]);
    real wsum = 0;
    for (int s = 0; s < 2; ++s) {
      int gb = s == 0 ? k->body_a : k->body_b;
      int b = gb < 0 ? -1 : gb - W->b0;
      S->cb[2 * c + s] = b;
      real* Jb = S->J + (2 * c + s) * 18;
      if (b < 0) { memset(Jb, 0, sizeof(real) * 18); continue; }
      real sg = s == 0 ? -1 : 1;
      v3 r = sub(p, V(S->xc[3 * b], S->xc[3 * b + 1], S->xc[3 * b + 2]));
      for (int row = 0; row < 3; ++row) {
        v3 ang = cross(r, F[row]);
        Jb[6 * row + 0] = sg * F[row].x; Jb[6 * row + 1] = sg * F[row].y; Jb[6 * row + 2] = sg * F[row].z;
        Jb[6 * row + 3] = sg * ang.x; Jb[6 * row + 4] = sg * ang.y; Jb[6 * row + 5] = sg * ang.z;
      }
      /* trace(W_bb) = 3/m + |r|^2 tr(Iw^-1) - r^T Iw^-1 r */
      const real* Ii = S->Iwi + 9 * b;
      real tr = Ii[0] + Ii[4] + Ii[8];
      v3 Ir = V(Ii[0] * r.x + Ii[1] * r.y + Ii[2] * r.z, Ii[3] * r.x + Ii[4] * r.y + Ii[5] * r.z,
                Ii[6] * r.x + Ii[7] * r.y + Ii[8] * r.z);
      wsum += 3 / S->m[b] + dot(r, r) * tr - dot(r, Ir);
    }
    real wdiag = wsum / 3;
    if (C->contact_law == 1) {
      /* MuJoCo rows: pos = dist - margin (the margins as SolverMuJoCo keeps them), inactive at
         pos >= 0; K, B from solref with the time constant floored at 2h (refsafe) */
      double solref[2], solimp[5], mu;
      int condim;
      or_mj_contact_params(M, k->shape_a, k->shape_b, solref, solimp, &mu, &condim);
      double pos = M->mj_zero_margins ? (double)k->dist : (double)k->phi0;
      double imp = or_mj_impedance(solimp, pos);
      double dmax = clampr(solimp[1], OR_MJ_MINIMP, OR_MJ_MAXIMP);
      double K, B;
      if (solref[0] > 0) {
        double tc = maxr(solref[0], 2.0 * (double)h), dr = solref[1];
        K = 1.0 / (dmax * dmax * tc * tc * dr * dr);
        B = 2.0 / (dmax * tc);
      } else {
        K = -solref[0] / (dmax * dmax);
        B = -solref[1] / dmax;
      }
      double w = 0;
      for (int s2 = 0; s2 < 2; ++s2) if (S->cb[2 * c + s2] >= 0) w += 1.0 / S->m[S->cb[2 * c + s2]];
      if (condim != 1 && condim != 3) condim = 3;
      double wr = condim == 3 ? w * (1 + mu * mu) * 2 * mu * mu / C->impratio : w;
      double R = maxr(OR_MJ_MINVAL, wr * (1 - imp) / imp);
      acc vt[3];
      contact_velocity(S, c, qd, vt);  /* J v at the step's start */
      S->D[c] = pos < 0 ? (real)(1.0 / R) : 0;
      for (int r = 0; r < 3; ++r) S->what[3 * c + r] = (real)((1.0 - (double)h * B) * vt[r]);
      S->what[3 * c + 2] -= (real)((double)h * K * imp * pos);
      S->mu[c] = condim == 1 ? 0 : (real)mu;
      S->Rn[c] = S->Rt[c] = 1; S->vhat[c] = 0;
      continue;
    }
    real Rn_soft = 1 / (h * k->k * (h + tau_d));
    real Rn_rigid = (real)(C->beta * C->beta / (4.0 * OR_PI * OR_PI)) * wdiag;
    S->Rn[c] = maxr(Rn_rigid, Rn_soft);
    S->Rt[c] = (real)C->sigma * wdiag;
    S->vhat[c] = -k->phi0 / (h + tau_d);
    S->mu[c] = k->mu;
  }
}

static void sap_alloc(sap_ctx* S, int nb, int nc) {
  int N = 6 * nb;
  int ncc = nc > 0 ? nc : 1;
  S->m = (real*)calloc(nb, sizeof(real)); S->Iw = (real*)calloc(9 * nb, sizeof(real));
  S->Iwi = (real*)calloc(9 * nb, sizeof(real)); S->xc = (real*)calloc(3 * nb, sizeof(real));
  S->vstar = (acc*)calloc(N, sizeof(acc)); S->Ad = (acc*)calloc(N, sizeof(acc));
  S->cb = (int*)calloc(2 * ncc, sizeof(int)); S->J = (real*)calloc(36 * ncc, sizeof(real));
  S->Rt = (real*)calloc(ncc, sizeof(real)); S->Rn = (real*)calloc(ncc, sizeof(real));
  S->vhat = (real*)calloc(ncc, sizeof(real)); S->mu = (real*)calloc(ncc, sizeof(real));
  S->H = (real*)calloc((size_t)N * N, sizeof(real)); S->g = (acc*)calloc(N, sizeof(acc));
  S->dv = (acc*)calloc(N, sizeof(acc)); S->v = (acc*)calloc(N, sizeof(acc));
  S->gam = (acc*)calloc(3 * ncc, sizeof(acc)); S->G = (real*)calloc(9 * ncc, sizeof(real));
  S->tmp = (acc*)calloc(4 * N, sizeof(acc)); S->rtmp = (real*)calloc(N > 0 ? N : 1, sizeof(real));
  S->Adense = NULL; S->Jd = NULL;
  S->law = 0;
  S->D = (real*)calloc(ncc, sizeof(real)); S->what = (real*)calloc(3 * ncc, sizeof(real));
}
static void sap_free(sap_ctx* S) {
  free(S->m); free(S->Iw); free(S->Iwi); free(S->xc); free(S->vstar); free(S->Ad); free(S->cb); free(S->J);
  free(S->Rt); free(S->Rn); free(S->vhat); free(S->mu); free(S->H); free(S->g); free(S->dv); free(S->v);
  free(S->gam); free(S->G); free(S->tmp); free(S->rtmp); free(S->D); free(S->what);
}

/* Semi-implicit update (solver.py:50-94 without damping): x_com += h v,
 * q = normalize(q + 1/2 h (w, 0) q), origin = x_com - R com.  Always in double.
 * mj_quat: the rotation as MuJoCo integrates a free joint (mju_quatIntegrate), by the exact
 * exponential q = exp(h w / 2) q of the new angular velocity, normalised. */
static void integrate_world_q(const world_view* W, const double* q, const acc* v, double h, double* q_out,
                              double* qd_out, int mj_quat) {
  const or_model* M = W->M;
  for (int b = 0; b < W->nb; ++b) {
    int gb = W->b0 + b;
    const double* bq = q + 7 * b;
    double qx = bq[3], qy = bq[4], qz = bq[5], qw = bq[6];
    double cx = M->body_com[3 * gb], cy = M->body_com[3 * gb + 1], cz = M->body_com[3 * gb + 2];
    /* COM position */
    double R0[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
                    2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
                    2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
    double xc[3];
    for (int i = 0; i < 3; ++i) xc[i] = bq[i] + R0[3 * i] * cx + R0[3 * i + 1] * cy + R0[3 * i + 2] * cz;
    const acc* vb = v + 6 * b;
    for (int i = 0; i < 3; ++i) xc[i] += h * (double)vb[i];
    double wx = vb[3], wy = vb[4], wz = vb[5];
    double nx, ny, nz, nw;
    if (mj_quat) {
      double wn = sqrt(wx * wx + wy * wy + wz * wz), th = 0.5 * h * wn;
      double sc = wn > 0 ? sin(th) / wn : 0.0, cw = cos(th);
      double ex = wx * sc, ey = wy * sc, ez = wz * sc;  /* exp(h w / 2) = (e, cw), times q */
      nx = cw * qx + ex * qw + ey * qz - ez * qy;
      ny = cw * qy + ey * qw + ez * qx - ex * qz;
      nz = cw * qz + ez * qw + ex * qy - ey * qx;
      nw = cw * qw - ex * qx - ey * qy - ez * qz;
    } else {
      /* dq = (w, 0) * q */
      double dx = wy * qz - wz * qy + wx * qw;
      double dy = wz * qx - wx * qz + wy * qw;
      double dz = wx * qy - wy * qx + wz * qw;
      double dw = -(wx * qx + wy * qy + wz * qz);
      nx = qx + 0.5 * h * dx; ny = qy + 0.5 * h * dy; nz = qz + 0.5 * h * dz; nw = qw + 0.5 * h * dw;
    }
    double l = sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
    nx /= l; ny /= l; nz /= l; nw /= l;
    double R1[9] = {1 - 2 * (ny * ny + nz * nz), 2 * (nx * ny - nz * nw), 2 * (nx * nz + ny * nw),
                    2 * (nx * ny + nz * nw), 1 - 2 * (nx * nx + nz * nz), 2 * (ny * nz - nx * nw),
                    2 * (nx * nz - ny * nw), 2 * (ny * nz + nx * nw), 1 - 2 * (nx * nx + ny * ny)};
    double* o = q_out + 7 * b;
    for (int i = 0; i < 3; ++i) o[i] = xc[i] - (R1[3 * i] * cx + R1[3 * i + 1] * cy + R1[3 * i + 2] * cz);
    o[3] = nx; o[4] = ny; o[5] = nz; o[6] = nw;
    for (int i = 0; i < 6; ++i) qd_out[6 * b + i] = vb[i];
  }
}

/* One inner step (= CENIC ComputeNextContinuousState, solver_sap_adaptive.py:714-729):
 * given contacts computed at the step's start pose, solve and integrate.
 * Returns 1 when the convex solve converged. */
static int inner_step(const world_view* W, const or_config* C, const double* q, const double* qd, const double* f,
                      double h, const or_contact* cs, int nc, const acc* guess, double* q_out, double* qd_out,
                      acc* v_out, solve_stats* st) {
  sap_ctx S;
  sap_alloc(&S, W->nb, nc);
  sap_setup(&S, W, C, q, qd, f, (real)h, cs, nc);
  acc* v = S.v;
  for (int i = 0; i < S.N; ++i) v[i] = guess ? guess[i] : (acc)qd[i];
  int ok = sap_newton(&S, C, v, st);
  integrate_world_q(W, q, v, h, q_out, qd_out, C->contact_law == 1);
  if (v_out) memcpy(v_out, v, sizeof(acc) * S.N);
  sap_free(&S);
  return ok;
}

/* Exported single-step entry: contacts generated at q, one solve + integrate of world w.
 * q/qd/f/q_out/qd_out hold that world's bodies only. */
int or_substep_world(const or_model* M, const or_config* C, int w, const double* q, const double* qd,
                     const double* f, double h, const double* guess, double* q_out, double* qd_out, int* stats) {
  world_view W = {M, w, M->world_body_start[w], M->world_body_start[w + 1] - M->world_body_start[w]};
  int cap = C->max_contacts;
  or_contact* cs = (or_contact*)malloc(sizeof(or_contact) * (cap > 0 ? cap : 1));
  int np;
  int nc = or_collide_world(M, w, q, cs, cap, &np);
  if (nc > cap) nc = cap;
  solve_stats st;
  acc* g = NULL;
  if (guess) {
    g = (acc*)malloc(sizeof(acc) * 6 * (W.nb ? W.nb : 1));
    for (int i = 0; i < 6 * W.nb; ++i) g[i] = guess[i];
  }
  int ok = inner_step(&W, C, q, qd, f, h, cs, nc, g, q_out, qd_out, NULL, &st);
  free(g);
  if (stats) { stats[0] = st.iterations; stats[1] = st.factorizations; stats[2] = st.converged; stats[3] = nc; }
  free(cs);
  return ok;
}

/* Exported convex problem of one inner step of world w (contacts at q) and its solution, for
   checks that re-evaluate the SAP optimality from the problem data alone: N = 6 nb dofs, A [N,N]
   (dense), v* [N], contact rows J [3 nc, N] (t1, t2, n per contact), Rt, Rn, v_hat, mu [nc], and
   the converged velocity v [N].  Capacities: N <= n_cap, nc <= c_cap.  Returns nc (or -1 when a
   capacity is too small); *converged = the solve met its optimality test. */
int or_substep_problem(const or_model* M, const or_config* C, int w, const double* q, const double* qd,
                       const double* f, double h, int n_cap, int c_cap, double* A_out, double* vstar_out,
                       double* J_out, double* Rt_out, double* Rn_out, double* vhat_out, double* mu_out,
                       double* v_out, int* converged, double* D_out, double* what_out) {
  world_view W = {M, w, M->world_body_start[w], M->world_body_start[w + 1] - M->world_body_start[w]};
  int cap = C->max_contacts;
  or_contact* cs = (or_contact*)malloc(sizeof(or_contact) * (cap > 0 ? cap : 1));
  int np;
  int nc = or_collide_world(M, w, q, cs, cap, &np);
  if (nc > cap) nc = cap;
  int N = 6 * W.nb;
  if (N > n_cap || nc > c_cap) { free(cs); return -1; }
  sap_ctx S;
  sap_alloc(&S, W.nb, nc);
  sap_setup(&S, &W, C, q, qd, f, (real)h, cs, nc);
  acc* v = S.v;
  for (int i = 0; i < N; ++i) v[i] = qd[i];
  solve_stats st;
  *converged = sap_newton(&S, C, v, &st);
  memset(A_out, 0, sizeof(double) * N * N);
  acc* e = S.tmp + 3 * N;  /* unit vectors through apply_A: the dense A */
  for (int j = 0; j < N; ++j) {
    acc col[6 * 64];
    for (int i = 0; i < N; ++i) e[i] = (acc)(i == j);
    apply_A(&S, e, col);
    for (int i = 0; i < N; ++i) A_out[i * N + j] = col[i];
  }
  memset(J_out, 0, sizeof(double) * 3 * nc * N);
  for (int c = 0; c < nc; ++c) {
    for (int s2 = 0; s2 < 2; ++s2) {
      int b = S.cb[2 * c + s2];
      if (b < 0) continue;
      const real* Jb = S.J + (2 * c + s2) * 18;
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 6; ++k) J_out[(3 * c + r) * N + 6 * b + k] += Jb[6 * r + k];
    }
    Rt_out[c] = S.Rt[c]; Rn_out[c] = S.Rn[c]; vhat_out[c] = S.vhat[c]; mu_out[c] = S.mu[c];
    /* contact law 1: the rows' D and target contact velocity */
    D_out[c] = S.D[c];
    for (int r = 0; r < 3; ++r) what_out[3 * c + r] = S.what[3 * c + r];
  }
  for (int i = 0; i < N; ++i) { vstar_out[i] = S.vstar[i]; v_out[i] = v[i]; }
  sap_free(&S);
  free(cs);
  return nc;
}

/* ====================================================================
 * Step-doubling attempt and frame loop (solver_sap_adaptive.py:732-960,
 * solver_mujoco_adaptive.py:692-849), FP64 controller.
 * ==================================================================== */
typedef struct {
  double *dt, *dt_half, *ideal_dt, *sim_time, *next_time, *last_error, *accepted_error;
  uint8_t *accepted, *commit, *diverged;
  int32_t *substeps, *iterations, *solve_fail;
  int64_t* cum_accepted;
  double* trend;       /* [B*6] mean acceleration of each body over its last accepted step (warm_start 1, 2;
                          NULL: zero) */
} or_ctrl;

/* Contact records of the device's parity entry points (16 values per contact: shape a, shape b,
   body a, body b (global, -1 static), manifold slot, unit normal A->B, world midpoint, distance,
   phi0, stiffness, friction, 0) as solver contacts; the writer radii and margins are not used by
   the solve and are left zero. */
static void contacts_from_records(const double* rec, int n, or_contact* out) {
  for (int i = 0; i < n; ++i) {
    const double* r = rec + 16 * i;
    or_contact* c = out + i;
    memset(c, 0, sizeof(*c));
    c->shape_a = (int)r[0]; c->shape_b = (int)r[1]; c->body_a = (int)r[2]; c->body_b = (int)r[3];
    c->sub = (int)r[4];
    for (int k = 0; k < 3; ++k) { c->n[k] = (real)r[5 + k]; c->c[k] = (real)r[8 + k]; }
    c->dist = (real)r[11]; c->phi0 = (real)r[12]; c->k = (real)r[13]; c->mu = (real)r[14];
  }
}

/* Attempt of one world from (q, qd): writes full/double results, returns error (sentinel on NaN).
   given (NULL: collide) holds the contact sets to solve with at q and at the midpoint state,
   given[0..given_n[0]) and given[cap..cap + given_n[1]) with cap = C->max_contacts.
   Newton seeds (solver_sap_adaptive.py:756-771 for warm_start 0): the full step from qd, the
   first half step from (qd + v_full) / 2, the second half step from v_full.  warm_start 1 seeds
   the full step at qd + dt/2 trend (trend: the world's [nb*6] mean accelerations over its last
   accepted step, NULL = zero) and the second half step at 2 v_mid - qd; warm_start 2 also seeds
   the first half step at qd + dt/4 trend instead of from v_full; the problems solved are the
   same. */
static double attempt(const world_view* W, const or_config* C, const double* q, const double* qd, const double* f,
                      const double* trend, double dt, double* q_full, double* qd_full, double* q_dbl,
                      double* qd_dbl, int* solve_ok, int* n_contacts, const or_contact* given,
                      const int32_t* given_n, int* n_contacts_mid) {
  int nb = W->nb, N = 6 * nb;
  int cap = C->max_contacts;
  or_contact* cs = (or_contact*)malloc(sizeof(or_contact) * (cap > 0 ? cap : 1));
  acc* vfull = (acc*)malloc(sizeof(acc) * (N > 0 ? N : 1));
  acc* guess = (acc*)malloc(sizeof(acc) * (N > 0 ? N : 1));
  double* q_mid = (double*)malloc(sizeof(double) * 7 * (nb > 0 ? nb : 1));
  double* qd_mid = (double*)malloc(sizeof(double) * 6 * (nb > 0 ? nb : 1));
  solve_stats st;
  int ok = 1, np, nc;
  if (given) {
    nc = given_n[0] < cap ? given_n[0] : cap;
    memcpy(cs, given, sizeof(or_contact) * nc);
  } else {
    nc = or_collide_world(W->M, W->w, q, cs, cap, &np);
  }
  if (nc > cap) nc = cap;
  if (n_contacts) *n_contacts = nc;
  for (int i = 0; i < N; ++i) guess[i] = C->warm_start >= 1 && trend ? qd[i] + 0.5 * dt * trend[i] : qd[i];
  ok &= inner_step(W, C, q, qd, f, dt, cs, nc, guess, q_full, qd_full, vfull, &st);
  double err = 0.0;
  if (C->mode == 1) {
    for (int i = 0; i < N; ++i)
      guess[i] = C->warm_start == 2 ? (trend ? qd[i] + 0.25 * dt * trend[i] : qd[i]) : 0.5 * (qd[i] + vfull[i]);
    ok &= inner_step(W, C, q, qd, f, 0.5 * dt, cs, nc, guess, q_mid, qd_mid, NULL, &st);
    if (given) {
      nc = given_n[1] < cap ? given_n[1] : cap;
      memcpy(cs, given + cap, sizeof(or_contact) * nc);
    } else {
      nc = or_collide_world(W->M, W->w, q_mid, cs, cap, &np);
    }
    if (nc > cap) nc = cap;
    if (n_contacts_mid) *n_contacts_mid = nc;
    if (C->warm_start >= 1)
      for (int i = 0; i < N; ++i) vfull[i] = 2.0 * qd_mid[i] - qd[i];  /* the second half step's seed */
    ok &= inner_step(W, C, q_mid, qd_mid, f, 0.5 * dt, cs, nc, vfull, q_dbl, qd_dbl, NULL, &st);
    for (int i = 0; i < 7 * nb; ++i) {
      double d = fabs(q_dbl[i] - q_full[i]);
      if (isnan(d)) { err = NAN; break; }
      if (d > err) err = d;
    }
  } else {
    memcpy(q_dbl, q_full, sizeof(double) * 7 * nb);
    memcpy(qd_dbl, qd_full, sizeof(double) * 6 * nb);
    for (int i = 0; i < 7 * nb; ++i) if (isnan(q_full[i])) err = NAN;
  }
  if (isnan(err) || isinf(err)) err = 1.0e10;
  *solve_ok = ok;
  free(cs); free(vfull); free(guess); free(q_mid); free(qd_mid);
  return err;
}

/* Exported single attempt for parity tests: out_err[w], and full/double states of every world. */
void or_attempt_all(const or_model* M, const or_config* C, const double* q, const double* qd, const double* f,
                    const double* trend, const double* dt, double* q_full, double* qd_full, double* q_dbl,
                    double* qd_dbl, double* err, int32_t* solve_ok, int32_t* n_contacts, int32_t* n_contacts_mid) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int w = 0; w < M->world_count; ++w) {
    int b0 = M->world_body_start[w];
    world_view W = {M, w, b0, M->world_body_start[w + 1] - b0};
    int ok, nc, ncm = 0;
    err[w] = attempt(&W, C, q + 7 * b0, qd + 6 * b0, f + 6 * b0, trend ? trend + 6 * b0 : NULL, dt[w],
                     q_full + 7 * b0, qd_full + 6 * b0, q_dbl + 7 * b0, qd_dbl + 6 * b0, &ok, &nc, NULL, NULL, &ncm);
    solve_ok[w] = ok;
    n_contacts[w] = nc;
    if (n_contacts_mid) n_contacts_mid[w] = ncm;
  }
}

/* Exported attempt of every world with given contact sets (records [W][2][C->max_contacts][16],
   counts [W][2]: the contact sets a device attempt solved with at q and at its midpoint): the
   solves, integration and error norm of the attempt on the same contacts, in the layout of
   or_attempt_all. */
void or_attempt_all_given(const or_model* M, const or_config* C, const double* q, const double* qd, const double* f,
                          const double* trend, const double* dt, const double* records, const int32_t* counts,
                          double* q_full, double* qd_full, double* q_dbl, double* qd_dbl, double* err,
                          int32_t* solve_ok) {
  int cap = C->max_contacts > 0 ? C->max_contacts : 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int w = 0; w < M->world_count; ++w) {
    int b0 = M->world_body_start[w];
    world_view W = {M, w, b0, M->world_body_start[w + 1] - b0};
    or_contact* given = (or_contact*)malloc(sizeof(or_contact) * 2 * cap);
    const int32_t* cw = counts + 2 * w;
    int32_t n[2] = {cw[0] < cap ? cw[0] : cap, cw[1] < cap ? cw[1] : cap};
    const double* rw = records + (size_t)w * 2 * cap * 16;
    contacts_from_records(rw, n[0], given);
    contacts_from_records(rw + 16 * cap, n[1], given + cap);
    int ok, nc;
    err[w] = attempt(&W, C, q + 7 * b0, qd + 6 * b0, f + 6 * b0, trend ? trend + 6 * b0 : NULL, dt[w],
                     q_full + 7 * b0, qd_full + 6 * b0, q_dbl + 7 * b0, qd_dbl + 6 * b0, &ok, &nc, given, n, NULL);
    solve_ok[w] = ok;
    free(given);
  }
}

/* Whole boundary call for every world: open frame, march attempts until the boundary.
 * Per-world marching is exactly the lock-step loop of the reference (done worlds are no-ops).
 * q/qd are updated in place.  Returns 0 ok, 2 if any inner solve failed (RuntimeError in
 * the reference, solver_sap_adaptive.py:900-905). */
/* Controller after one attempt of world w (size `step`, error `err`, ok = every solve converged):
   _apply_solve_convergence_to_error + _adapt_dt (solver_sap_adaptive.py:780-808) or
   _calc_adjusted_step + the MuJoCo commit rules (solver_mujoco_adaptive.py:835-905).  Returns
   whether to commit; adds accepted steps to *cum. */
static int ctrl_after_attempt(const or_config* C, or_ctrl* K, int w, double err, int ok, double step,
                              double eff_dt_max, int64_t* cum) {
  if (!ok) { err = C->divergence_threshold; K->solve_fail[w] = 1; }
  K->last_error[w] = err;
  int commit;
  if (C->controller == 0) {
    int32_t cum1 = 0;
    or_adapt_dt_f64(1, &K->last_error[w], &K->sim_time[w], &K->next_time[w], &K->dt[w], &K->dt_half[w],
                    &K->ideal_dt[w], &K->diverged[w], &K->accepted[w], &K->accepted_error[w], &K->substeps[w],
                    &cum1, C->mode, C->tol, C->dt_min, eff_dt_max, C->divergence_threshold);
    *cum += cum1;
    commit = K->accepted[w];
  } else {
    or_calc_adjusted_step_f64(1, &K->last_error[w], &K->dt[w], &K->ideal_dt[w], &K->accepted[w], &K->commit[w],
                              &K->diverged[w], C->tol, C->dt_min, C->divergence_threshold);
    commit = K->commit[w];
    if (K->accepted[w]) {
      K->sim_time[w] += K->dt[w];
      K->accepted_error[w] = err;
      if (step > 0) { K->substeps[w] += 1; *cum += 1; }
    }
    if (K->diverged[w]) K->sim_time[w] = K->next_time[w];
    double d = K->ideal_dt[w];
    d = d < C->dt_min ? C->dt_min : (d > eff_dt_max ? eff_dt_max : d);
    K->dt[w] = d; K->dt_half[w] = 0.5 * d;
  }
  return commit;
}

/* Frame opening of world w: _open_frame + _seed_dt (solver_sap_adaptive.py:89-132) or _apply_dt_cap,
   _rebase_time, _boundary_advance (solver_mujoco_adaptive.py:805-833). */
static void ctrl_open_frame(const or_config* C, or_ctrl* K, int w, double dt_outer, double eff_dt_max) {
  K->iterations[w] = 0;
  if (C->controller == 0) {
    double base = K->next_time[w];
    K->sim_time[w] = K->sim_time[w] - base;
    K->next_time[w] = dt_outer;
    double d = C->mode == 0 ? C->dt_init : K->ideal_dt[w];
    d = d < C->dt_min ? C->dt_min : (d > eff_dt_max ? eff_dt_max : d);
    K->dt[w] = d; K->dt_half[w] = 0.5 * d;
    K->substeps[w] = 0; K->diverged[w] = 0;
  } else {
    double d = K->ideal_dt[w];
    d = d < C->dt_min ? C->dt_min : (d > eff_dt_max ? eff_dt_max : d);
    K->dt[w] = d; K->dt_half[w] = 0.5 * d;
    double base = K->next_time[w];
    K->sim_time[w] -= base; K->next_time[w] -= base;
    K->next_time[w] += dt_outer;
    K->diverged[w] = 0;
  }
}

/* _clamp_dt_to_boundary for world w; returns the attempt size */
static double ctrl_clamp(or_ctrl* K, int w) {
  double rem = K->next_time[w] - K->sim_time[w];
  if (rem <= 0) { K->dt[w] = 0; K->dt_half[w] = 0; }
  else if (K->dt[w] > rem) { K->dt[w] = rem; K->dt_half[w] = 0.5 * rem; }
  return K->dt[w];
}

int or_step_frame(const or_model* M, const or_config* C, double* q, double* qd, const double* f, or_ctrl* K,
                  double dt_outer) {
  int status = 0;
  double eff_dt_max = C->dt_max < dt_outer ? C->dt_max : dt_outer;
  int64_t cum = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : status) reduction(+ : cum)
  for (int w = 0; w < M->world_count; ++w) {
    int b0 = M->world_body_start[w], nb = M->world_body_start[w + 1] - b0;
    world_view W = {M, w, b0, nb};
    double *qw = q + 7 * b0, *qdw = qd + 6 * b0;
    const double* fw = f + 6 * b0;
    double* qf = (double*)malloc(sizeof(double) * 7 * (nb ? nb : 1));
    double* qdf = (double*)malloc(sizeof(double) * 6 * (nb ? nb : 1));
    double* qdb = (double*)malloc(sizeof(double) * 7 * (nb ? nb : 1));
    double* qddb = (double*)malloc(sizeof(double) * 6 * (nb ? nb : 1));
    ctrl_open_frame(C, K, w, dt_outer, eff_dt_max);
    for (int it = 0; it < C->max_substeps; ++it) {
      if (!(K->sim_time[w] < K->next_time[w])) break;
      K->iterations[w]++;
      double step = ctrl_clamp(K, w);
      int ok = 1, nc;
      double err;
      double* tw = K->trend && C->warm_start >= 1 ? K->trend + 6 * b0 : NULL;
      if (step > 0) err = attempt(&W, C, qw, qdw, fw, tw, step, qf, qdf, qdb, qddb, &ok, &nc, NULL, NULL, NULL);
      else { err = 0; memcpy(qdb, qw, sizeof(double) * 7 * nb); memcpy(qddb, qdw, sizeof(double) * 6 * nb);
             memcpy(qf, qw, sizeof(double) * 7 * nb); memcpy(qdf, qdw, sizeof(double) * 6 * nb); }
      if (!ok) status = 2;
      int commit = ctrl_after_attempt(C, K, w, err, ok, step, eff_dt_max, &cum);
      if (commit) {
        const double* sq = C->mode == 0 ? qf : qdb;
        const double* sqd = C->mode == 0 ? qdf : qddb;
        if (tw && K->accepted[w] && step > 0)
          for (int i = 0; i < 6 * nb; ++i) {
            double a = (sqd[i] - qdw[i]) / step;
            tw[i] = isfinite(a) ? a : 0.0;
          }
        memcpy(qw, sq, sizeof(double) * 7 * nb);
        memcpy(qdw, sqd, sizeof(double) * 6 * nb);
      }
      if (!ok) break;
    }
    free(qf); free(qdf); free(qdb); free(qddb);
  }
  K->cum_accepted[0] += cum;
  return status;
}

/* ====================================================================
 * Articulated worlds: generalized-coordinate step doubling.
 *
 * Joints are REVOLUTE (1 dof) or FREE (root only, parent = world, identity joint frames), listed
 * parent-before-child; joint j of world w moves body world_body_start[w] + (j - world_joint_start[w]).
 * Coordinates: REVOLUTE angle / rate; FREE (origin position, quaternion xyzw) / (COM velocity,
 * angular velocity), all world frame -- Newton's joint_q / joint_qd conventions (state.py:149-153).
 * Spatial vectors are world-frame Pluecker coordinates at the world origin, (angular; linear).
 * Dynamics: M(q) = sum_b J_b^T I_b J_b, bias c(q, v) from Newton-Euler (gravity included),
 * implicit PD drives tau = ke (q_t - q - h v') - kd v' on revolute dofs, armature on the diagonal:
 *   A = M + diag(armature + h (kd + h ke)),  A v* = (M + diag(armature)) v + h (ke (q_t - q) - c).
 * The convex contact problem is the free-body one with dense A and contact rows J_c = F^T (P_b J_b
 * - P_a J_a), P = [-[p]x, I] the point-velocity map; R_n, R_t use tr(J_c A^-1 J_c^T) / 3.
 * ==================================================================== */
#define OR_JT_REVOLUTE 1
#define OR_JT_FREE 4
#define OR_ART_W_MIN 1.0e-10 /* Delassus diagonal below which a contact cannot move [1/kg] */
typedef struct or_art {
  int joint_count;
  const int32_t* world_joint_start; /* [W+1] */
  const int32_t* joint_type;
  const int32_t* joint_parent;      /* global body index, -1 world */
  const double* joint_X_p;          /* [J,7] joint frame in the parent body (or world) frame */
  const double* joint_X_c;          /* [J,7] joint frame in the child body frame */
  const double* joint_axis;         /* [J,3] revolute axis in the joint frame */
  const int32_t* joint_q_start;     /* [J+1] */
  const int32_t* joint_qd_start;    /* [J+1] */
  const double* target_pos;         /* [dofs] */
  const double* target_ke;
  const double* target_kd;
  const double* armature;
  /* control inputs (control.py:31-110), NULL = none: velocity targets and generalized forces per
     dof (a FREE joint's joint_f is a COM wrench, world frame), feedforward per dof, applied COM
     wrenches per body [B,6] (world frame, State.body_f) */
  const double* target_vel;
  const double* joint_f;
  const double* joint_act;
  const double* body_f;
} or_art;

typedef struct { double p[3]; double R[9]; } rigid;  /* world pose, R row-major */

static void rigid_from7(const double* x, rigid* T) {
  qt q = {x[3], x[4], x[5], x[6]};
  m33 R = qmat(q);
  for (int i = 0; i < 9; ++i) T->R[i] = R.m[i];
  for (int i = 0; i < 3; ++i) T->p[i] = x[i];
}
static void rigid_mul(const rigid* a, const rigid* b, rigid* o) {
  rigid t;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j)
      t.R[3 * i + j] = a->R[3 * i] * b->R[j] + a->R[3 * i + 1] * b->R[3 + j] + a->R[3 * i + 2] * b->R[6 + j];
    t.p[i] = a->p[i] + a->R[3 * i] * b->p[0] + a->R[3 * i + 1] * b->p[1] + a->R[3 * i + 2] * b->p[2];
  }
  *o = t;
}
static void rigid_inv(const rigid* a, rigid* o) {
  rigid t;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) t.R[3 * i + j] = a->R[3 * j + i];
  for (int i = 0; i < 3; ++i) t.p[i] = -(t.R[3 * i] * a->p[0] + t.R[3 * i + 1] * a->p[1] + t.R[3 * i + 2] * a->p[2]);
  *o = t;
}
static void rot_axis_angle(const double* a, double th, rigid* o) {
  double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  double x = a[0] / n, y = a[1] / n, z = a[2] / n, s = sin(0.5 * th), c = cos(0.5 * th);
  double x7[7] = {0, 0, 0, x * s, y * s, z * s, c};
  rigid_from7(x7, o);
}
/* rotation matrix -> quaternion (x, y, z, w), w >= 0 branch-stable */
static void rot_to_quat(const double* R, double* q) {
  double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = sqrt(tr + 1.0) * 2;
    q[3] = 0.25 * s; q[0] = (R[7] - R[5]) / s; q[1] = (R[2] - R[6]) / s; q[2] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    q[3] = (R[7] - R[5]) / s; q[0] = 0.25 * s; q[1] = (R[1] + R[3]) / s; q[2] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    q[3] = (R[2] - R[6]) / s; q[0] = (R[1] + R[3]) / s; q[1] = 0.25 * s; q[2] = (R[5] + R[7]) / s;
  } else {
    double s = sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    q[3] = (R[3] - R[1]) / s; q[0] = (R[2] + R[6]) / s; q[1] = (R[5] + R[7]) / s; q[2] = 0.25 * s;
  }
}

typedef struct {
  int j0, nj, b0, nb, q0, nq, d0, nv;
  rigid* X;       /* [nb] body poses */
  double* S;      /* [nv, 6] motion subspace columns */
  int* dof_body;  /* [nv] body moved by the dof's joint */
  int* parent;    /* [nb] local parent body, -1 world */
} art_kin;

static void art_alloc(const or_model* M, const or_art* A, int w, art_kin* K) {
  K->j0 = A->world_joint_start[w];
  K->nj = A->world_joint_start[w + 1] - K->j0;
  K->b0 = M->world_body_start[w];
  K->nb = M->world_body_start[w + 1] - K->b0;
  K->q0 = A->joint_q_start[K->j0];
  K->nq = A->joint_q_start[K->j0 + K->nj] - K->q0;
  K->d0 = A->joint_qd_start[K->j0];
  K->nv = A->joint_qd_start[K->j0 + K->nj] - K->d0;
  K->X = (rigid*)calloc(K->nb > 0 ? K->nb : 1, sizeof(rigid));
  K->S = (double*)calloc(6 * (K->nv > 0 ? K->nv : 1), sizeof(double));
  K->dof_body = (int*)calloc(K->nv > 0 ? K->nv : 1, sizeof(int));
  K->parent = (int*)calloc(K->nb > 0 ? K->nb : 1, sizeof(int));
}
static void art_free(art_kin* K) { free(K->X); free(K->S); free(K->dof_body); free(K->parent); }

/* forward kinematics and motion subspaces of world w at its coordinates q [nq] */
static void art_fk(const or_model* M, const or_art* A, art_kin* K, const double* q) {
  for (int jl = 0; jl < K->nj; ++jl) {
    int j = K->j0 + jl, b = jl;
    int pg = A->joint_parent[j];
    int p = pg < 0 ? -1 : pg - K->b0;
    K->parent[b] = p;
    const double* qj = q + (A->joint_q_start[j] - K->q0);
    int dj = A->joint_qd_start[j] - K->d0;
    if (A->joint_type[j] == OR_JT_FREE) {
      rigid_from7(qj, &K->X[b]);
      const double* com = M->body_com + 3 * (K->b0 + b);
      double c[3];
      for (int i = 0; i < 3; ++i)
        c[i] = K->X[b].p[i] + K->X[b].R[3 * i] * com[0] + K->X[b].R[3 * i + 1] * com[1] + K->X[b].R[3 * i + 2] * com[2];
      /* dofs (v_com, w): spatial (w; v_O) with v_O = v_com - w x c = v_com + c x w */
      for (int k = 0; k < 3; ++k) {
        double* s = K->S + 6 * (dj + k);
        for (int i = 0; i < 6; ++i) s[i] = 0;
        s[3 + k] = 1;
        double* t = K->S + 6 * (dj + 3 + k);
        double e[3] = {0, 0, 0};
        e[k] = 1;
        for (int i = 0; i < 3; ++i) t[i] = e[i];
        t[3] = c[1] * e[2] - c[2] * e[1]; t[4] = c[2] * e[0] - c[0] * e[2]; t[5] = c[0] * e[1] - c[1] * e[0];
      }
      for (int k = 0; k < 6; ++k) K->dof_body[dj + k] = b;
    } else {
      rigid Xp, Xj, Xc, Xci, R;
      rigid_from7(A->joint_X_p + 7 * j, &Xp);
      if (p >= 0) rigid_mul(&K->X[p], &Xp, &Xj);
      else Xj = Xp;
      rot_axis_angle(A->joint_axis + 3 * j, qj[0], &R);
      rigid_from7(A->joint_X_c + 7 * j, &Xc);
      rigid_inv(&Xc, &Xci);
      rigid tmp;
      rigid_mul(&Xj, &R, &tmp);
      rigid_mul(&tmp, &Xci, &K->X[b]);
      const double* ax = A->joint_axis + 3 * j;
      double an = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
      double a[3], o[3];
      for (int i = 0; i < 3; ++i) a[i] = (Xj.R[3 * i] * ax[0] + Xj.R[3 * i + 1] * ax[1] + Xj.R[3 * i + 2] * ax[2]) / an;
      for (int i = 0; i < 3; ++i) o[i] = Xj.p[i];
      double* s = K->S + 6 * dj;
      s[0] = a[0]; s[1] = a[1]; s[2] = a[2];
      s[3] = o[1] * a[2] - o[2] * a[1]; s[4] = o[2] * a[0] - o[0] * a[2]; s[5] = o[0] * a[1] - o[1] * a[0];
      K->dof_body[dj] = b;
    }
  }
}
/* does body b's path from the root contain the joint of body a (a is b or an ancestor)? */
static int art_on_path(const art_kin* K, int a, int b) {
  for (int x = b; x >= 0; x = K->parent[x]) if (x == a) return 1;
  return 0;
}
/* spatial inertia of body b at the world origin, (angular; linear) ordering, 6x6 row-major */
static void art_inertia(const or_model* M, const art_kin* K, int b, double* I6, double* c_out) {
  int gb = K->b0 + b;
  const rigid* X = &K->X[b];
  const double* com = M->body_com + 3 * gb;
  const double* Ib = M->body_inertia + 9 * gb;
  double m = M->body_mass[gb], c[3], Ic[9];
  for (int i = 0; i < 3; ++i) c[i] = X->p[i] + X->R[3 * i] * com[0] + X->R[3 * i + 1] * com[1] + X->R[3 * i + 2] * com[2];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
    double a = 0;
    for (int k = 0; k < 3; ++k) for (int l = 0; l < 3; ++l) a += X->R[3 * i + k] * Ib[3 * k + l] * X->R[3 * j + l];
    Ic[3 * i + j] = a;
  }
  double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  double cx[9] = {0, -c[2], c[1], c[2], 0, -c[0], -c[1], c[0], 0};
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
    I6[6 * i + j] = Ic[3 * i + j] + m * ((i == j ? cc : 0) - c[i] * c[j]);
    I6[6 * i + 3 + j] = m * cx[3 * i + j];
    I6[6 * (3 + i) + j] = -m * cx[3 * i + j];
    I6[6 * (3 + i) + 3 + j] = i == j ? m : 0;
  }
  if (c_out) for (int i = 0; i < 3; ++i) c_out[i] = c[i];
}
static void mcross(const double* V, const double* m, double* o) { /* motion cross V x m */
  const double *w = V, *v = V + 3, *a = m, *b = m + 3;
  o[0] = w[1] * a[2] - w[2] * a[1]; o[1] = w[2] * a[0] - w[0] * a[2]; o[2] = w[0] * a[1] - w[1] * a[0];
  o[3] = w[1] * b[2] - w[2] * b[1] + v[1] * a[2] - v[2] * a[1];
  o[4] = w[2] * b[0] - w[0] * b[2] + v[2] * a[0] - v[0] * a[2];
  o[5] = w[0] * b[1] - w[1] * b[0] + v[0] * a[1] - v[1] * a[0];
}
static void fcross(const double* V, const double* f, double* o) { /* force cross V x* f */
  const double *w = V, *v = V + 3, *n = f, *fl = f + 3;
  o[0] = w[1] * n[2] - w[2] * n[1] + v[1] * fl[2] - v[2] * fl[1];
  o[1] = w[2] * n[0] - w[0] * n[2] + v[2] * fl[0] - v[0] * fl[2];
  o[2] = w[0] * n[1] - w[1] * n[0] + v[0] * fl[1] - v[1] * fl[0];
  o[3] = w[1] * fl[2] - w[2] * fl[1]; o[4] = w[2] * fl[0] - w[0] * fl[2]; o[5] = w[0] * fl[1] - w[1] * fl[0];
}

/* Mass matrix M [nv,nv] and bias c [nv] (velocity products and gravity) at (q, v). */
static void art_dynamics(const or_model* M, const or_art* A, const art_kin* K, const double* v, double* Mm,
                         double* c) {
  int nv = K->nv, nb = K->nb;
  memset(Mm, 0, sizeof(double) * nv * nv);
  memset(c, 0, sizeof(double) * nv);
  double* Vb = (double*)calloc(6 * (nb ? nb : 1), sizeof(double));
  double* ab = (double*)calloc(6 * (nb ? nb : 1), sizeof(double));
  /* body velocities V_b = sum_{dofs on path} S_k v_k */
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < nv; ++k)
      if (art_on_path(K, K->dof_body[k], b))
        for (int i = 0; i < 6; ++i) Vb[6 * b + i] += K->S[6 * k + i] * v[k];
  /* velocity-product accelerations: revolute Sdot = V_child x S; FREE (v_com, w) dofs: the
     angular columns carry c x e_k, whose rate gives (0; v_com x w) */
  double* sdq = (double*)calloc(6 * (nv ? nv : 1), sizeof(double));
  for (int jl = 0; jl < K->nj; ++jl) {
    int j = K->j0 + jl, dj = A->joint_qd_start[j] - K->d0, b = jl;
    if (A->joint_type[j] == OR_JT_FREE) {
      const double *vc = v + dj, *w = v + dj + 3;
      double* o = sdq + 6 * dj;
      o[3] = vc[1] * w[2] - vc[2] * w[1]; o[4] = vc[2] * w[0] - vc[0] * w[2]; o[5] = vc[0] * w[1] - vc[1] * w[0];
    } else {
      double t[6];
      mcross(Vb + 6 * b, K->S + 6 * dj, t);
      for (int i = 0; i < 6; ++i) sdq[6 * dj + i] = t[i] * v[dj];
    }
  }
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < nv; ++k)
      if (art_on_path(K, K->dof_body[k], b))
        for (int i = 0; i < 6; ++i) ab[6 * b + i] += sdq[6 * k + i];
  for (int b = 0; b < nb; ++b) {
    double I6[36], cb[3], f[6], IV[6], t[6];
    art_inertia(M, K, b, I6, cb);
    double m = M->body_mass[K->b0 + b];
    for (int i = 0; i < 6; ++i) {
      f[i] = 0; IV[i] = 0;
      for (int k = 0; k < 6; ++k) { f[i] += I6[6 * i + k] * ab[6 * b + k]; IV[i] += I6[6 * i + k] * Vb[6 * b + k]; }
    }
    fcross(Vb + 6 * b, IV, t);
    /* gravity plus the applied COM wrench (fe; te): body_f of the body, and joint_f of a FREE
       joint that moves it, both world frame; (c x fe + te; fe) at the origin */
    const double* g = M->gravity;
    double fe[3] = {m * g[0], m * g[1], m * g[2]}, te[3] = {0, 0, 0};
    int gb = K->b0 + b, jj = K->j0 + b;
    if (A->body_f)
      for (int i = 0; i < 3; ++i) { fe[i] += A->body_f[6 * gb + i]; te[i] += A->body_f[6 * gb + 3 + i]; }
    if (A->joint_f && A->joint_type[jj] == OR_JT_FREE) {
      int d = A->joint_qd_start[jj];
      for (int i = 0; i < 3; ++i) { fe[i] += A->joint_f[d + i]; te[i] += A->joint_f[d + 3 + i]; }
    }
    double fg[6] = {cb[1] * fe[2] - cb[2] * fe[1] + te[0], cb[2] * fe[0] - cb[0] * fe[2] + te[1],
                    cb[0] * fe[1] - cb[1] * fe[0] + te[2], fe[0], fe[1], fe[2]};
    for (int i = 0; i < 6; ++i) f[i] += t[i] - fg[i];
    /* J_b^T f and J_b^T I J_b */
    for (int k = 0; k < nv; ++k) {
      if (!art_on_path(K, K->dof_body[k], b)) continue;
      const double* sk = K->S + 6 * k;
      for (int i = 0; i < 6; ++i) c[k] += sk[i] * f[i];
      double Isk[6];
      for (int i = 0; i < 6; ++i) { Isk[i] = 0; for (int l = 0; l < 6; ++l) Isk[i] += I6[6 * i + l] * sk[l]; }
      for (int l = 0; l < nv; ++l) {
        if (!art_on_path(K, K->dof_body[l], b)) continue;
        const double* sl = K->S + 6 * l;
        double s = 0;
        for (int i = 0; i < 6; ++i) s += sl[i] * Isk[i];
        Mm[l * nv + k] += s;
      }
    }
  }
  free(Vb); free(ab); free(sdq);
}

/* body poses [nb,7] of world w at coordinates q (for collision) */
static void art_body_q(const art_kin* K, double* bq) {
  for (int b = 0; b < K->nb; ++b) {
    for (int i = 0; i < 3; ++i) bq[7 * b + i] = K->X[b].p[i];
    rot_to_quat(K->X[b].R, bq + 7 * b + 3);
  }
}

/* Integrate coordinates: REVOLUTE angle += h rate; FREE: COM += h v_com, quaternion += 1/2 h (w,0) q
   (normalised), origin = COM - R com (solver.py:50-94 without damping). */
static void art_integrate(const or_model* M, const or_art* A, const art_kin* K, const double* q, const double* v,
                          double h, double* qo) {
  memcpy(qo, q, sizeof(double) * K->nq);
  for (int jl = 0; jl < K->nj; ++jl) {
    int j = K->j0 + jl, qj = A->joint_q_start[j] - K->q0, dj = A->joint_qd_start[j] - K->d0;
    if (A->joint_type[j] == OR_JT_FREE) {
      const double* bq = q + qj;
      const double* com = M->body_com + 3 * (K->b0 + jl);
      double qx = bq[3], qy = bq[4], qz = bq[5], qw = bq[6];
      double R0[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
                      2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
                      2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
      double xc[3];
      for (int i = 0; i < 3; ++i) xc[i] = bq[i] + R0[3 * i] * com[0] + R0[3 * i + 1] * com[1] + R0[3 * i + 2] * com[2] + h * v[dj + i];
      double wx = v[dj + 3], wy = v[dj + 4], wz = v[dj + 5];
      double dx = wy * qz - wz * qy + wx * qw, dy = wz * qx - wx * qz + wy * qw;
      double dz = wx * qy - wy * qx + wz * qw, dw = -(wx * qx + wy * qy + wz * qz);
      double nx = qx + 0.5 * h * dx, ny = qy + 0.5 * h * dy, nz = qz + 0.5 * h * dz, nw = qw + 0.5 * h * dw;
      double l = sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
      nx /= l; ny /= l; nz /= l; nw /= l;
      double R1[9] = {1 - 2 * (ny * ny + nz * nz), 2 * (nx * ny - nz * nw), 2 * (nx * nz + ny * nw),
                      2 * (nx * ny + nz * nw), 1 - 2 * (nx * nx + nz * nz), 2 * (ny * nz - nx * nw),
                      2 * (nx * nz - ny * nw), 2 * (ny * nz + nx * nw), 1 - 2 * (nx * nx + ny * ny)};
      double* o = qo + qj;
      for (int i = 0; i < 3; ++i) o[i] = xc[i] - (R1[3 * i] * com[0] + R1[3 * i + 1] * com[1] + R1[3 * i + 2] * com[2]);
      o[3] = nx; o[4] = ny; o[5] = nz; o[6] = nw;
    } else {
      qo[qj] = q[qj] + h * v[dj];
    }
  }
}

/* One inner step of articulated world w (contacts cs at q): returns 1 when the solve converged. */
static int art_inner_step(const or_model* M, const or_art* A, const or_config* C, int w, const double* q,
                          const double* v0, double h, const or_contact* cs, int nc, const double* guess, double* q_out,
                          double* v_out, solve_stats* st) {
  art_kin K;
  art_alloc(M, A, w, &K);
  art_fk(M, A, &K, q);
  int nv = K.nv;
  double* Mm = (double*)calloc((size_t)nv * nv + 1, sizeof(double));
  double* cb = (double*)calloc(nv + 1, sizeof(double));
  art_dynamics(M, A, &K, v0, Mm, cb);
  sap_ctx S;
  sap_alloc(&S, 1, nc);
  /* sap_alloc sized its vectors for 6 dofs; rebuild them for nv */
  free(S.vstar); free(S.Ad); free(S.H); free(S.g); free(S.dv); free(S.v); free(S.tmp); free(S.rtmp);
  S.N = nv; S.nb = 0; S.nc = nc;
  S.vstar = (acc*)calloc(nv + 1, sizeof(acc)); S.Ad = (acc*)calloc(nv + 1, sizeof(acc));
  S.H = (real*)calloc((size_t)nv * nv + 1, sizeof(real)); S.g = (acc*)calloc(nv + 1, sizeof(acc));
  S.dv = (acc*)calloc(nv + 1, sizeof(acc)); S.v = (acc*)calloc(nv + 1, sizeof(acc));
  S.tmp = (acc*)calloc(4 * nv + 4, sizeof(acc)); S.rtmp = (real*)calloc(nv + 1, sizeof(real));
  S.Adense = (real*)calloc((size_t)nv * nv + 1, sizeof(real));
  S.Jd = (real*)calloc((size_t)3 * (nc > 0 ? nc : 1) * nv, sizeof(real));
  double* rhs = (double*)calloc(nv + 1, sizeof(double));
  for (int i = 0; i < nv; ++i) {
    double acc = 0;
    for (int k = 0; k < nv; ++k) acc += Mm[i * nv + k] * v0[k];
    rhs[i] = acc;
  }
  for (int i = 0; i < nv * nv; ++i) S.Adense[i] = (real)Mm[i];
  for (int jl = 0; jl < K.nj; ++jl) {
    int j = K.j0 + jl, dj = A->joint_qd_start[j] - K.d0, qj = A->joint_q_start[j] - K.q0;
    int nd = A->joint_type[j] == OR_JT_FREE ? 6 : 1;
    for (int k = 0; k < nd; ++k) {
      int gd = K.d0 + dj + k, i = dj + k;
      double arm = A->armature[gd], ke = A->target_ke[gd], kd = A->target_kd[gd];
      S.Adense[(size_t)i * nv + i] += (real)(arm + h * (kd + h * ke));
      rhs[i] += arm * v0[i];
      if (A->joint_type[j] == OR_JT_REVOLUTE) rhs[i] += h * ke * (A->target_pos[gd] - q[qj]);
      /* velocity target, feedforward, and the joint force of a revolute dof */
      if (A->target_vel) rhs[i] += h * kd * A->target_vel[gd];
      if (A->joint_act) rhs[i] += h * A->joint_act[gd];
      if (A->joint_f && A->joint_type[j] == OR_JT_REVOLUTE) rhs[i] += h * A->joint_f[gd];
    }
  }
  for (int i = 0; i < nv; ++i) rhs[i] -= h * cb[i];
  /* v* = A^-1 rhs */
  real* L = (real*)calloc((size_t)nv * nv + 1, sizeof(real));
  memcpy(L, S.Adense, sizeof(real) * nv * nv);
  int spd = cholesky(L, nv);
  for (int i = 0; i < nv; ++i) S.rtmp[i] = (real)rhs[i];
  if (spd) chol_solve(L, nv, S.rtmp);
  for (int i = 0; i < nv; ++i) S.vstar[i] = S.rtmp[i];
  for (int i = 0; i < nv; ++i) S.Ad[i] = S.Adense[(size_t)i * nv + i];
  /* contact rows and regularisation */
  real tau_d = (real)C->tau_d;
  real* y = (real*)calloc(nv + 1, sizeof(real));
  for (int c = 0; c < nc; ++c) {
    const or_contact* k = cs + c;
    v3 n = V(k->n[0], k->n[1], k->n[2]);
    v3 e = RA(n.x) < RA(n.y) ? (RA(n.x) < RA(n.z) ? V(1, 0, 0) : V(0, 0, 1)) : (RA(n.y) < RA(n.z) ? V(0, 1, 0) : V(0, 0, 1));
    v3 t1 = nrm(cross(n, e)), t2 = cross(n, t1);
    v3 F[3] = {t1, t2, n};
    double p[3] = {k->c[0], k->c[1], k->c[2]};
    real* Jc = S.Jd + (size_t)3 * c * nv;
    for (int s = 0; s < 2; ++s) {
      int gb = s == 0 ? k->body_a : k->body_b;
      if (gb < 0) continue;
      int b = gb - K.b0;
      double sg = s == 0 ? -1 : 1;
      for (int d = 0; d < nv; ++d) {
        if (!art_on_path(&K, K.dof_body[d], b)) continue;
        const double* sd = K.S + 6 * d;
        /* point velocity of column d: v_O + w x p */
        double pv[3] = {sd[3] + sd[1] * p[2] - sd[2] * p[1], sd[4] + sd[2] * p[0] - sd[0] * p[2],
                        sd[5] + sd[0] * p[1] - sd[1] * p[0]};
        for (int r = 0; r < 3; ++r)
          Jc[(size_t)r * nv + d] += (real)(sg * (F[r].x * pv[0] + F[r].y * pv[1] + F[r].z * pv[2]));
      }
    }
    /* w = tr(J A^-1 J^T) / 3 via L^-1 J^T */
    real wsum = 0;
    for (int r = 0; r < 3; ++r) {
      for (int i = 0; i < nv; ++i) {
        real s2 = Jc[(size_t)r * nv + i];
        for (int kk = 0; kk < i; ++kk) s2 -= L[(size_t)i * nv + kk] * y[kk];
        y[i] = s2 / L[(size_t)i * nv + i];
        wsum += y[i] * y[i];
      }
    }
    real wdiag = wsum / 3;
    if (wdiag < (real)OR_ART_W_MIN) {
      /* the contact point cannot move (e.g. on a joint axis of a fixed base): no constraint */
      for (size_t i = 0; i < (size_t)3 * nv; ++i) Jc[i] = 0;
      S.Rn[c] = 1; S.Rt[c] = 1; S.vhat[c] = 0; S.mu[c] = k->mu;
      continue;
    }
    real Rn_soft = 1 / ((real)h * k->k * ((real)h + tau_d));
    real Rn_rigid = (real)(C->beta * C->beta / (4.0 * OR_PI * OR_PI)) * wdiag;
    S.Rn[c] = maxr(Rn_rigid, Rn_soft);
    S.Rt[c] = (real)C->sigma * wdiag;
    S.vhat[c] = -k->phi0 / ((real)h + tau_d);
    S.mu[c] = k->mu;
  }
  acc* v = S.v;
  for (int i = 0; i < nv; ++i) v[i] = guess ? guess[i] : v0[i];
  int ok = spd && sap_newton(&S, C, v, st);
  if (!spd) { st->iterations = 0; st->factorizations = 0; st->converged = 0; }
  double* vd = (double*)calloc(nv + 1, sizeof(double));
  for (int i = 0; i < nv; ++i) vd[i] = v[i];
  art_integrate(M, A, &K, q, vd, h, q_out);
  memcpy(v_out, vd, sizeof(double) * nv);
  free(vd); free(y); free(L); free(rhs); free(Mm); free(cb); free(S.Adense); free(S.Jd);
  sap_free(&S);
  art_free(&K);
  return ok;
}

static int art_collide(const or_model* M, const or_art* A, const or_config* C, int w, const double* q,
                       or_contact* cs) {
  art_kin K;
  art_alloc(M, A, w, &K);
  art_fk(M, A, &K, q);
  double* bq = (double*)calloc(7 * (K.nb ? K.nb : 1), sizeof(double));
  art_body_q(&K, bq);
  int np;
  int nc = or_collide_world(M, w, bq, cs, C->max_contacts, &np);
  free(bq);
  art_free(&K);
  return nc < C->max_contacts ? nc : C->max_contacts;
}

/* Exported: body poses [nb,7] of world w at its coordinates q [nq] */
void or_art_fk(const or_model* M, const or_art* A, int w, const double* q, double* body_q) {
  art_kin K;
  art_alloc(M, A, w, &K);
  art_fk(M, A, &K, q);
  art_body_q(&K, body_q);
  art_free(&K);
}
/* Exported: mass matrix [nv,nv] and bias [nv] of world w (tests) */
void or_art_dynamics(const or_model* M, const or_art* A, int w, const double* q, const double* v, double* Mm,
                     double* c) {
  art_kin K;
  art_alloc(M, A, w, &K);
  art_fk(M, A, &K, q);
  art_dynamics(M, A, &K, v, Mm, c);
  art_free(&K);
}
/* Exported single inner step of world w: stats = (newton iterations, factorizations, converged, contacts) */
int or_art_substep(const or_model* M, const or_art* A, const or_config* C, int w, const double* q, const double* v,
                   double h, double* q_out, double* v_out, int* stats) {
  or_contact* cs = (or_contact*)malloc(sizeof(or_contact) * (C->max_contacts > 0 ? C->max_contacts : 1));
  int nc = art_collide(M, A, C, w, q, cs);
  solve_stats st;
  int ok = art_inner_step(M, A, C, w, q, v, h, cs, nc, NULL, q_out, v_out, &st);
  if (stats) { stats[0] = st.iterations; stats[1] = st.factorizations; stats[2] = st.converged; stats[3] = nc; }
  free(cs);
  return ok;
}
/* Step-doubling attempt of articulated world w from its coordinates q / velocities v (the world's
   slices) with step h: full / doubled results into the world's slices, FP64 max-norm error
   (1e10 sentinel), *ok = every solve converged, *n_contacts at q.  Fixed mode (C->mode 0) takes
   the full step only (the doubled result is the full one, error 0 unless non-finite). */
static double art_attempt_world(const or_model* M, const or_art* A, const or_config* C, int w, const double* q,
                                const double* v, double h, double* q_full, double* v_full, double* q_dbl,
                                double* v_dbl, int* ok_out, int* n_contacts, const or_contact* given,
                                const int32_t* given_n, int* n_contacts_mid) {
  int j0 = A->world_joint_start[w], j1 = A->world_joint_start[w + 1];
  int nq = A->joint_q_start[j1] - A->joint_q_start[j0], nv = A->joint_qd_start[j1] - A->joint_qd_start[j0];
  or_contact* cs = (or_contact*)malloc(sizeof(or_contact) * (C->max_contacts > 0 ? C->max_contacts : 1));
  double* qm = (double*)calloc(nq + 1, sizeof(double));
  double* vm = (double*)calloc(nv + 1, sizeof(double));
  double* g = (double*)calloc(nv + 1, sizeof(double));
  solve_stats st;
  int ok = 1;
  int cap = C->max_contacts > 0 ? C->max_contacts : 1;
  int nc;
  if (given) { nc = given_n[0] < cap ? given_n[0] : cap; memcpy(cs, given, sizeof(or_contact) * nc); }
  else nc = art_collide(M, A, C, w, q, cs);
  if (n_contacts) *n_contacts = nc;
  ok &= art_inner_step(M, A, C, w, q, v, h, cs, nc, NULL, q_full, v_full, &st);
  double e = 0;
  if (C->mode == 1) {
    for (int i = 0; i < nv; ++i) g[i] = 0.5 * (v[i] + v_full[i]);
    ok &= art_inner_step(M, A, C, w, q, v, 0.5 * h, cs, nc, g, qm, vm, &st);
    if (given) { nc = given_n[1] < cap ? given_n[1] : cap; memcpy(cs, given + cap, sizeof(or_contact) * nc); }
    else nc = art_collide(M, A, C, w, qm, cs);
    if (n_contacts_mid) *n_contacts_mid = nc;
    ok &= art_inner_step(M, A, C, w, qm, vm, 0.5 * h, cs, nc, v_full, q_dbl, v_dbl, &st);
    for (int i = 0; i < nq; ++i) {
      double d = fabs(q_dbl[i] - q_full[i]);
      if (isnan(d)) { e = NAN; break; }
      if (d > e) e = d;
    }
  } else {
    memcpy(q_dbl, q_full, sizeof(double) * nq);
    memcpy(v_dbl, v_full, sizeof(double) * nv);
    for (int i = 0; i < nq; ++i) if (isnan(q_full[i])) e = NAN;
  }
  if (isnan(e) || isinf(e)) e = 1.0e10;
  *ok_out = ok;
  free(cs); free(qm); free(vm); free(g);
  return e;
}

/* Exported step-doubling attempt of every articulated world from (q, v) with per-world dt:
   full / doubled coordinates and velocities, FP64 error norms (1e10 sentinel), solve status. */
void or_art_attempt_all(const or_model* M, const or_art* A, const or_config* C, const double* q, const double* v,
                        const double* dt, double* q_full, double* v_full, double* q_dbl, double* v_dbl, double* err,
                        int32_t* solve_ok, int32_t* n_contacts, int32_t* n_contacts_mid) {
  int W = M->world_count;
#pragma omp parallel for schedule(dynamic, 1)
  for (int w = 0; w < W; ++w) {
    int j0 = A->world_joint_start[w];
    int q0 = A->joint_q_start[j0], d0 = A->joint_qd_start[j0];
    int ok = 1, nc = 0, ncm = 0;
    err[w] = art_attempt_world(M, A, C, w, q + q0, v + d0, dt[w], q_full + q0, v_full + d0, q_dbl + q0, v_dbl + d0,
                               &ok, &nc, NULL, NULL, &ncm);
    solve_ok[w] = ok;
    n_contacts[w] = nc;
    if (n_contacts_mid) n_contacts_mid[w] = ncm;
  }
}

/* Exported attempt of every articulated world with given contact sets (records and counts as in
   or_attempt_all_given), in the layout of or_art_attempt_all. */
void or_art_attempt_all_given(const or_model* M, const or_art* A, const or_config* C, const double* q,
                              const double* v, const double* dt, const double* records, const int32_t* counts,
                              double* q_full, double* v_full, double* q_dbl, double* v_dbl, double* err,
                              int32_t* solve_ok) {
  int cap = C->max_contacts > 0 ? C->max_contacts : 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int w = 0; w < M->world_count; ++w) {
    int j0 = A->world_joint_start[w];
    int q0 = A->joint_q_start[j0], d0 = A->joint_qd_start[j0];
    or_contact* given = (or_contact*)malloc(sizeof(or_contact) * 2 * cap);
    const int32_t* cw = counts + 2 * w;
    int32_t n[2] = {cw[0] < cap ? cw[0] : cap, cw[1] < cap ? cw[1] : cap};
    const double* rw = records + (size_t)w * 2 * cap * 16;
    contacts_from_records(rw, n[0], given);
    contacts_from_records(rw + 16 * cap, n[1], given + cap);
    int ok, nc;
    err[w] = art_attempt_world(M, A, C, w, q + q0, v + d0, dt[w], q_full + q0, v_full + d0, q_dbl + q0, v_dbl + d0,
                               &ok, &nc, given, n, NULL);
    solve_ok[w] = ok;
    free(given);
  }
}

/* Exported boundary call of articulated worlds: every world advances by dt_outer in place over its
   coordinates q / velocities v with the controller state in K (the frame loop of or_step_frame
   with the articulated attempt).  Returns 2 when a convex solve failed. */
int or_art_step_frame(const or_model* M, const or_art* A, const or_config* C, double* q, double* v, or_ctrl* K,
                      double dt_outer) {
  int status = 0;
  double eff_dt_max = C->dt_max < dt_outer ? C->dt_max : dt_outer;
  int64_t cum = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : status) reduction(+ : cum)
  for (int w = 0; w < M->world_count; ++w) {
    int j0 = A->world_joint_start[w], j1 = A->world_joint_start[w + 1];
    int q0 = A->joint_q_start[j0], nq = A->joint_q_start[j1] - q0;
    int d0 = A->joint_qd_start[j0], nv = A->joint_qd_start[j1] - d0;
    double *qw = q + q0, *vw = v + d0;
    double* qf = (double*)calloc(nq + 1, sizeof(double));
    double* vf = (double*)calloc(nv + 1, sizeof(double));
    double* qb = (double*)calloc(nq + 1, sizeof(double));
    double* vb = (double*)calloc(nv + 1, sizeof(double));
    ctrl_open_frame(C, K, w, dt_outer, eff_dt_max);
    for (int it = 0; it < C->max_substeps; ++it) {
      if (!(K->sim_time[w] < K->next_time[w])) break;
      K->iterations[w]++;
      double step = ctrl_clamp(K, w);
      int ok = 1, nc = 0;
      double err = 0;
      if (step > 0 && j1 > j0)
        err = art_attempt_world(M, A, C, w, qw, vw, step, qf, vf, qb, vb, &ok, &nc, NULL, NULL, NULL);
      else { memcpy(qf, qw, sizeof(double) * nq); memcpy(qb, qw, sizeof(double) * nq);
             memcpy(vf, vw, sizeof(double) * nv); memcpy(vb, vw, sizeof(double) * nv); }
      if (!ok) status = 2;
      int commit = ctrl_after_attempt(C, K, w, err, ok, step, eff_dt_max, &cum);
      if (commit) {
        memcpy(qw, C->mode == 0 ? qf : qb, sizeof(double) * nq);
        memcpy(vw, C->mode == 0 ? vf : vb, sizeof(double) * nv);
      }
      if (!ok) break;
    }
    free(qf); free(vf); free(qb); free(vb);
  }
  K->cum_accepted[0] += cum;
  return status;
}

/* ---------------------------------------------------------------- MuJoCo coordinate adapters */
/* FREE joint with the world as parent frame (kernels.py:705-789 Newton -> MuJoCo, :578-666 back):
   qpos = (origin, quaternion wxyz); qvel = (origin velocity in world, angular velocity in the body
   frame); Newton joint_q = (origin, quaternion xyzw), joint_qd = (COM velocity, angular velocity),
   both world. */
void or_coords_to_mujoco(int B, const double* com, const double* jq, const double* jqd, double* qpos,
                         double* qvel) {
  for (int b = 0; b < B; ++b) {
    const double* q = jq + 7 * b;
    const double* u = jqd + 6 * b;
    qt r = {q[3], q[4], q[5], q[6]};
    m33 R = qmat(r);
    v3 c = mv(&R, V(com[3 * b], com[3 * b + 1], com[3 * b + 2]));
    v3 w = V(u[3], u[4], u[5]);
    v3 vo = sub(V(u[0], u[1], u[2]), cross(w, c));
    v3 wb = mtv(&R, w);
    double* p = qpos + 7 * b;
    p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[6]; p[4] = q[3]; p[5] = q[4]; p[6] = q[5];
    double* v = qvel + 6 * b;
    v[0] = vo.x; v[1] = vo.y; v[2] = vo.z; v[3] = wb.x; v[4] = wb.y; v[5] = wb.z;
  }
}
void or_coords_from_mujoco(int B, const double* com, const double* qpos, const double* qvel, double* jq,
                           double* jqd) {
  for (int b = 0; b < B; ++b) {
    const double* p = qpos + 7 * b;
    const double* v = qvel + 6 * b;
    double* q = jq + 7 * b;
    q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = p[4]; q[4] = p[5]; q[5] = p[6]; q[6] = p[3];
    qt r = {q[3], q[4], q[5], q[6]};
    m33 R = qmat(r);
    v3 w = mv(&R, V(v[3], v[4], v[5]));
    v3 c = mv(&R, V(com[3 * b], com[3 * b + 1], com[3 * b + 2]));
    v3 vc = add(V(v[0], v[1], v[2]), cross(w, c));
    double* u = jqd + 6 * b;
    u[0] = vc.x; u[1] = vc.y; u[2] = vc.z; u[3] = w.x; u[4] = w.y; u[5] = w.z;
  }
}

/* OpenMP threads of the following parallel regions (the CPU baseline times all cores and one) */
void or_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
