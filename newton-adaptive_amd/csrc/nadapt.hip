// nadapt: gfx950 kernels and the C ABI of the batched step-doubling rigid-body step.
//
// One wavefront owns one world.  The served boundary call (na_step) is a single persistent
// launch: waves pull worlds from an atomic queue and march each world to its boundary
// entirely on the device (collide at q_t, full solve, half solve, collide at q_mid, half
// solve, FP64 error norm, FP64 controller, masked commit), keeping the world's pose (FP64),
// velocity and controller scalars in registers between attempts.  The lock-step form
// (na_step_plain) runs the same attempt as one launch per iteration over all worlds with a
// 4-byte flag read back per iteration, like the reference's host loop
// (solver_sap_adaptive.py:877-907).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nadapt.h"
#include "na_art.h"

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;
static int fail(const char* msg) {
  g_err = msg;
  return -1;
}
#define HIPCHK(x)                                                               \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      g_err = std::string(#x) + ": " + hipGetErrorString(e_);                   \
      return -1;                                                                \
    }                                                                           \
  } while (0)

extern "C" const char* na_last_error_string(void) { return g_err.c_str(); }
// error reporting for the other translation units of the library
int na_set_error(const char* msg) { return fail(msg); }

// ---------------------------------------------------------------- objects
struct na_model {
  NaModel M;
  int max_nb, max_seg, max_nl;
  bool articulated;
  bool has_mj;  // the MuJoCo shape arrays are present (contact law 1)
};

struct Ctrl {
  double *dt, *dt_half, *ideal_dt, *sim_time, *next_time, *last_error, *acc_error;
  uint8_t *accepted, *commit, *diverged;
  int32_t *substeps, *iterations, *solve_fail, *overflow;
  unsigned long long* cum;
  long long* work;  // [W, NA_WORK]
  float* trend;     // [B, 6] each body's mean acceleration over its last committed step (warm_start 1, 2)
};
// per-world work counters: attempts, Newton iterations, CG iterations, contacts, pairs, solves,
// then shader-clock cycles of collision, solve setup, residual evaluation, preconditioner, PCG,
// line search (the cycle counters stay zero unless built with -DNA_PHASE_TIMERS)
#define NA_WORK 12

struct na_solver {
  const na_model* model;
  NaConfig cfg;
  Ctrl K;
  double* q64;
  int32_t* queue;  // [2]
  int32_t* flag;   // [1]
  int32_t* flag_host;
  const int32_t* order;
  int grid_waves;
  int tiles;        // leading entries of a served launch's list run as two-wave tiles (0: none)
  int plain_batch;  // lock-step form: finished worlds still run each iteration's collision passes
  void* pool;
};

// ---------------------------------------------------------------- controller (FP64 on device)
#define DRAKE_SAFETY 0.9
#define DRAKE_MIN_SHRINK 0.1
#define DRAKE_MAX_GROW 5.0
#define DRAKE_HYST_HIGH 1.2
#define DRAKE_HYST_LOW 0.9

template <typename T>
__device__ __forceinline__ T drake_new_step(T step, T e, T tol) {
  T ee = e > (T)1.0e-30 ? e : (T)1.0e-30;
  T ns = (T)DRAKE_SAFETY * step * sqrt(tol / ee);
  if (ns > (T)DRAKE_HYST_LOW * step && ns < (T)DRAKE_HYST_HIGH * step) ns = step;
  T lo = (T)DRAKE_MIN_SHRINK * step, hi = (T)DRAKE_MAX_GROW * step;
  return ns < lo ? lo : (ns > hi ? hi : ns);
}
template <typename T>
__device__ __forceinline__ bool is_div_err(T e, T thr) {
  return isnan(e) || isinf(e) || e >= thr;
}

// _adapt_dt (solver_sap_adaptive.py:237-345); returns accept.  `acc_inc` counts accepted steps.
template <typename T>
__device__ __forceinline__ bool adapt_dt(T e, T& sim_time, T next_time, T& dt, T& dt_half, T& ideal_dt,
                                         uint8_t& diverged, T& acc_err, int& substeps, int& acc_inc, int mode,
                                         T tol, T dt_min, T dt_max, T thr) {
  T step = dt;
  if (sim_time >= next_time || step <= (T)0) return false;
  bool div = is_div_err(e, thr);
  if (mode == 0) {
    if (div) {
      sim_time = next_time;
      diverged = 1;
      return false;
    }
    sim_time = sim_time + step;
    acc_err = e;
    substeps += 1;
    acc_inc += 1;
    return true;
  }
  if (step <= dt_min * (T)1.001) {
    if (div) {
      sim_time = next_time;
      diverged = 1;
      ideal_dt = dt_min;
      return false;
    }
    sim_time = sim_time + step;
    acc_err = e;
    substeps += 1;
    acc_inc += 1;
    ideal_dt = drake_new_step(step, e, tol);
    return true;
  }
  if (div) {
    T ns = (T)DRAKE_MIN_SHRINK * step;
    ideal_dt = ns;
    T d = ns < dt_min ? dt_min : (ns > dt_max ? dt_max : ns);
    dt = d;
    dt_half = d * (T)0.5;
    return false;
  }
  T ns = drake_new_step(step, e, tol);
  bool acc = (e <= tol) || (ns >= step);
  T d = ns < dt_min ? dt_min : (ns > dt_max ? dt_max : ns);
  if (acc) {
    sim_time = sim_time + step;
    acc_err = e;
    substeps += 1;
    acc_inc += 1;
  }
  ideal_dt = ns;
  dt = d;
  dt_half = d * (T)0.5;
  return acc;
}

// _calc_adjusted_step (solver_mujoco_adaptive.py:120-198)
template <typename T>
__device__ __forceinline__ void calc_adjusted_step(T e, T step, T& ideal_dt, uint8_t& accepted, uint8_t& commit,
                                                   uint8_t& diverged, T tol, T dt_min, T thr) {
  bool div = is_div_err(e, thr);
  if (step <= (T)0) {
    accepted = 1;
    commit = 1;
    return;
  }
  if (step <= dt_min * (T)1.001) {
    if (div) {
      accepted = 1;
      commit = 0;
      diverged = 1;
      ideal_dt = dt_min;
      return;
    }
    if (e > tol) {
      accepted = 1;
      commit = 1;
      ideal_dt = dt_min;
      return;
    }
  }
  if (div) {
    accepted = 0;
    commit = 0;
    ideal_dt = (T)DRAKE_MIN_SHRINK * step;
    return;
  }
  T ee = e > (T)1.0e-30 ? e : (T)1.0e-30;
  T ns = (T)DRAKE_SAFETY * step * sqrt(tol / ee);
  if (ns > (T)DRAKE_HYST_LOW * step && ns < (T)DRAKE_HYST_HIGH * step) ns = step;
  T lo = (T)DRAKE_MIN_SHRINK * step, hi = (T)DRAKE_MAX_GROW * step;
  ns = ns < lo ? lo : (ns > hi ? hi : ns);
  bool acc = (e <= tol) || (ns >= step);
  accepted = acc;
  commit = acc;
  ideal_dt = ns;
}

// ---------------------------------------------------------------- reference-semantics f32 kernels
__global__ void k_calc_adjusted_step_f32(int n, const float* err, const float* dt, float* ideal, uint8_t* acc,
                                         uint8_t* com, uint8_t* div, float tol, float dt_min, float thr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float idl = ideal[i];
  uint8_t a = acc[i], c = com[i], d = div[i];
  calc_adjusted_step<float>(err[i], dt[i], idl, a, c, d, tol, dt_min, thr);
  ideal[i] = idl;
  acc[i] = a;
  com[i] = c;
  div[i] = d;
}
__global__ void k_adapt_dt_f32(int n, const float* err, float* sim, const float* nxt, float* dt, float* dth,
                               float* ideal, uint8_t* div, uint8_t* acc, float* acc_err, int* sub, int* cum, int mode,
                               float tol, float dt_min, float dt_max, float thr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = sim[i], d = dt[i], dh = dth[i], idl = ideal[i], ae = acc_err[i];
  uint8_t dv = div[i];
  int sb = sub[i], inc = 0;
  bool a = adapt_dt<float>(err[i], s, nxt[i], d, dh, idl, dv, ae, sb, inc, mode, tol, dt_min, dt_max, thr);
  sim[i] = s; dt[i] = d; dth[i] = dh; ideal[i] = idl; acc_err[i] = ae; div[i] = dv; sub[i] = sb; acc[i] = a;
  if (inc) atomicAdd(cum, inc);
}
__global__ void k_rebase_time_f32(int n, float* sim, float* nxt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float b = nxt[i];
  sim[i] = sim[i] - b;
  nxt[i] = nxt[i] - b;
}
__global__ void k_reset_worlds_f32(int n, const uint8_t* mask, float dt_init, float* ideal, float* dt, float* dth,
                                   float* sim, float* nxt, uint8_t* div, uint8_t* acc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !mask[i]) return;
  ideal[i] = dt_init; dt[i] = dt_init; dth[i] = dt_init * 0.5f;
  sim[i] = 0.f; nxt[i] = 0.f; div[i] = 0; acc[i] = 0;
}
__global__ void k_clamp_dt_f32(int n, float* dt, float* dth, const float* sim, const float* nxt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float rem = nxt[i] - sim[i];
  if (rem <= 0.f) { dt[i] = 0.f; dth[i] = 0.f; }
  else if (dt[i] > rem) { dt[i] = rem; dth[i] = rem * 0.5f; }
}
__global__ void k_inf_norm_f32(int n, int coords, const float* qf, const float* qd, const float* scale, float* err) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  float m = 0.f;
  bool nan = false;
  for (int i = 0; i < coords; ++i) {
    float s = scale[w * coords + i] * fabsf(qd[w * coords + i] - qf[w * coords + i]);
    if (s != s) nan = true;
    m = fmaxf(m, s);
  }
  if (nan || isinf(m)) m = 1.0e10f;
  err[w] = m;
}

// ---------------------------------------------------------------- MuJoCo coordinate adapters
// Every body carries one FREE joint whose parent frame is the world, so joint and body rows
// coincide: Newton joint_q = (origin position, quaternion xyzw), joint_qd = (COM linear velocity,
// angular velocity), both world frame; MuJoCo qpos = (origin position, quaternion wxyz), qvel =
// (origin linear velocity in world, angular velocity in the body frame)
// (kernels.py:705-789 convert_warp_coords_to_mj_kernel, :578-666 the inverse).
__global__ void k_coords_to_mujoco(int B, const float* com, const float* jq, const float* jqd, float* qpos,
                                   float* qvel) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float* q = jq + 7 * b;
  const float* u = jqd + 6 * b;
  m3 R = qmat(q[3], q[4], q[5], q[6]);
  f3 c = mv(R, mk(com[3 * b], com[3 * b + 1], com[3 * b + 2]));
  f3 w = mk(u[3], u[4], u[5]);
  f3 vo = mk(u[0], u[1], u[2]) - cross3(w, c);
  f3 wb = mtv(R, w);
  float* p = qpos + 7 * b;
  p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
  p[3] = q[6]; p[4] = q[3]; p[5] = q[4]; p[6] = q[5];
  float* v = qvel + 6 * b;
  v[0] = vo.x; v[1] = vo.y; v[2] = vo.z; v[3] = wb.x; v[4] = wb.y; v[5] = wb.z;
}
__global__ void k_coords_from_mujoco(int B, const float* com, const float* qpos, const float* qvel, float* jq,
                                     float* jqd) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float* p = qpos + 7 * b;
  const float* v = qvel + 6 * b;
  float* q = jq + 7 * b;
  q[0] = p[0]; q[1] = p[1]; q[2] = p[2];
  q[3] = p[4]; q[4] = p[5]; q[5] = p[6]; q[6] = p[3];
  m3 R = qmat(q[3], q[4], q[5], q[6]);
  f3 w = mv(R, mk(v[3], v[4], v[5]));
  f3 c = mv(R, mk(com[3 * b], com[3 * b + 1], com[3 * b + 2]));
  f3 vc = mk(v[0], v[1], v[2]) + cross3(w, c);
  float* u = jqd + 6 * b;
  u[0] = vc.x; u[1] = vc.y; u[2] = vc.z; u[3] = w.x; u[4] = w.y; u[5] = w.z;
}

// Articulated models: one joint per body, joint j moves body j.  FREE roots convert as above
// after rotating the parent-frame twist by the joint's parent frame; revolute coordinates pass
// through (no reference offset).  qpos / qvel mirror the joint_q / joint_qd layout.
__global__ void k_coords_to_mujoco_art(int B, const int32_t* jtype, const int32_t* qs, const int32_t* ds,
                                       const float* Xp, const float* com, const float* jq, const float* jqd,
                                       float* qpos, float* qvel) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= B) return;
  const int q0 = qs[j], d0 = ds[j];
  if (jtype[j] != NA_JT_FREE) {
    qpos[q0] = jq[q0];
    qvel[d0] = jqd[d0];
    return;
  }
  const float* q = jq + q0;
  const float* u = jqd + d0;
  m3 Rp = qmat(Xp[7 * j + 3], Xp[7 * j + 4], Xp[7 * j + 5], Xp[7 * j + 6]);
  m3 R = qmat(q[3], q[4], q[5], q[6]);
  f3 vcw = mv(Rp, mk(u[0], u[1], u[2]));
  f3 w = mv(Rp, mk(u[3], u[4], u[5]));
  f3 c = mv(R, mk(com[3 * j], com[3 * j + 1], com[3 * j + 2]));
  f3 vo = vcw - cross3(w, c);
  f3 wb = mtv(R, w);
  float* p = qpos + q0;
  p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
  p[3] = q[6]; p[4] = q[3]; p[5] = q[4]; p[6] = q[5];
  float* v = qvel + d0;
  v[0] = vo.x; v[1] = vo.y; v[2] = vo.z; v[3] = wb.x; v[4] = wb.y; v[5] = wb.z;
}
__global__ void k_coords_from_mujoco_art(int B, const int32_t* jtype, const int32_t* qs, const int32_t* ds,
                                         const float* Xp, const float* com, const float* qpos, const float* qvel,
                                         float* jq, float* jqd) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= B) return;
  const int q0 = qs[j], d0 = ds[j];
  if (jtype[j] != NA_JT_FREE) {
    jq[q0] = qpos[q0];
    jqd[d0] = qvel[d0];
    return;
  }
  const float* p = qpos + q0;
  const float* v = qvel + d0;
  float* q = jq + q0;
  q[0] = p[0]; q[1] = p[1]; q[2] = p[2];
  q[3] = p[4]; q[4] = p[5]; q[5] = p[6]; q[6] = p[3];
  m3 Rp = qmat(Xp[7 * j + 3], Xp[7 * j + 4], Xp[7 * j + 5], Xp[7 * j + 6]);
  m3 R = qmat(q[3], q[4], q[5], q[6]);
  f3 w = mv(R, mk(v[3], v[4], v[5]));
  f3 c = mv(R, mk(com[3 * j], com[3 * j + 1], com[3 * j + 2]));
  f3 vc = mk(v[0], v[1], v[2]) + cross3(w, c);
  f3 vcp = mtv(Rp, vc), wp = mtv(Rp, w);
  float* u = jqd + d0;
  u[0] = vcp.x; u[1] = vcp.y; u[2] = vcp.z; u[3] = wp.x; u[4] = wp.y; u[5] = wp.z;
}

// ---------------------------------------------------------------- world-level device pipeline
struct WorldCtx {
  int w, b0, nb, nl, nseg;
};
__device__ __forceinline__ WorldCtx world_ctx(const NaModel& M, int w) {
  WorldCtx c;
  c.w = w;
  c.b0 = M.world_body_start[w];
  c.nb = M.world_body_end[w] - c.b0;
  c.nl = M.world_shape_end[w] - M.world_shape_start[w];
  c.nseg = c.nl + M.G;
  return c;
}

// Contacts of the world at this lane's body pose q (f64).  Returns the contact count kept.
// world_collide is the one out-of-line copy the step's forms share; the collision API's kernel
// inlines the body.
__device__ __forceinline__ int world_collide_body(const NaModel& M, const NaConfig& cfg, const WorldCtx& wc,
                                                  WorldLDS& S, const double* q, int& overflow, long long* wk) {
  const int l = lane_id();
  NA_TIC(t0);
  if (l < wc.nb) {
#pragma unroll
    for (int k = 0; k < 7; ++k) S.body.q[k][l] = (float)q[k];
  }
  wave_sync();
  stage_aabb(M, wc.w, wc.nseg, wc.nl, wc.b0, &S.body.q[0][0], NA_MAX_BODIES, S.seg);
  int np = stage_broad(M, wc.w, wc.nseg, wc.nl, cfg.broad_phase, S.seg);
  int cap = min(cfg.max_contacts, NA_MAX_CONTACTS);
  int nc = stage_narrow(M, wc.b0, np, cap, S.seg, S.con, cfg.contact_law == 1 ? &S.mj : nullptr);
  if (np > NA_MAX_PAIRS || nc > cap) overflow = 1;
  wk[3] += min(nc, cap);
  wk[4] += np;
  NA_TOC(wk[6], t0);
  return min(nc, cap);
}
__device__ __noinline__ int world_collide(const NaModel& M, const NaConfig& cfg, const WorldCtx& wc,
                                          lds_ptr_t<WorldLDS> Sp, const double* q, int& overflow, long long* wk) {
  return world_collide_body(M, cfg, wc, lds_deref<WorldLDS>(Sp), q, overflow, wk);
}

// One convex step of size h from (q, v0) with the contacts in LDS; integrates into qo/vo.
// The Newton iterate starts at guess64 when given (FP64), else at guess; v64, when given,
// receives the converged FP64 velocity.
__device__ bool world_solve(const NaModel& M, const NaConfig& cfg, const WorldCtx& wc, lds_ptr_t<WorldLDS> Sp, int nc,
                            const double* q, const float* v0, const float* f, double h, const float* guess, double* qo,
                            float* vo, long long* wk, SolveStats* out_st = nullptr, const double* guess64 = nullptr,
                            double* v64 = nullptr) {
  WorldLDS& S = lds_deref<WorldLDS>(Sp);
  const int l = lane_id();
  const bool own = l < wc.nb;
  NA_TIC(t0);
  BodyReg br;
  const bool mj = cfg.contact_law == 1;
  if (own) {
    body_setup(M, wc.b0 + l, q, v0, f, h, br, S.body, l);
    if (mj) {  // the start velocity, for the MuJoCo rows' reference velocity
#pragma unroll
      for (int k = 0; k < 6; ++k) S.body.v[k][l] = v0[k];
    }
  }
  if (mj && l == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) S.body.v[k][NA_DUMMY] = 0.f;
  }
  wave_sync();
  if (mj) contact_setup_mj(nc, (float)h, cfg, M.mj_zero_margins, S.con, S.mj, S.body, S.sol);
  else contact_setup(nc, (float)h, cfg, S.con, S.body, S.sol);
  build_entries(wc.nb, nc, S.con, S.sol);
  NA_TOC(wk[7], t0);
  double v[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = guess64 ? guess64[k] : (double)guess[k];
  SolveStats st;
  bool ok = sap_solve(wc.nb, nc, cfg, S.con, S.body, S.sol, br, v, st);
  if (own) integrate_body(M, wc.b0 + l, q, v, h, qo, mj);
#pragma unroll
  for (int k = 0; k < 6; ++k) vo[k] = (float)v[k];
  if (v64) {
#pragma unroll
    for (int k = 0; k < 6; ++k) v64[k] = v[k];
  }
  wk[1] += st.iterations;
  wk[2] += st.cg_iterations;
  wk[5] += 1;
  wk[8] += st.t_eval;
  wk[9] += st.t_pre;
  wk[10] += st.t_cg;
  wk[11] += st.t_ls;
  if (out_st) *out_st = st;
  return ok;
}

// Step-doubling attempt (solver_sap_adaptive.py:732-849): returns the FP64 error norm
// (1e10 sentinel for NaN/inf), `ok` false when a convex solve did not converge.  DUMP builds
// (parity entry point only) also write the contact sets of both collision passes to
// dump[2][NA_MAX_CONTACTS][NA_DUMP_FLOATS] and their counts to dump_n[2].
// Newton seeds of the three solves.  warm_start 0 is the reference's (solver_sap_adaptive.py:
// 756-771): the full step from v_t, the first half step from (v_t + v_full) / 2, the second from
// v_full.  warm_start 1 extrapolates: the full step from v_t + dt/2 a, with a = tr (this lane's
// mean acceleration over the last committed step; v_t when tr is null), the first half step from
// (v_t + v_full) / 2, the second from 2 v_mid - v_t (v_mid: the first half step's velocity).
// warm_start 2 ("concurrent") seeds the first half step at v_t + dt/4 a, independent of the full
// step's result, so the two solves that start at q_t need not wait for each other; the other two
// seeds are 1's.  The seeds change where Newton starts, not the problem each solve converges on.
template <bool DUMP = false>
__device__ double world_attempt(const NaModel& M, const NaConfig& cfg, const WorldCtx& wc, WorldLDS& S,
                                const double* q, const float* v, const float* f, const float* tr, double dt, double* qf,
                                float* vf, double* qd, float* vd, bool& ok, int& overflow, int& nc_out, long long* wk,
                                float* dump = nullptr, int32_t* dump_n = nullptr) {
  const int l = lane_id();
  const bool own = l < wc.nb;
  int nc = world_collide(M, cfg, wc, lds_ptr(S), q, overflow, wk);
  nc_out = nc;
  if (DUMP) {
    dump_contacts(S.seg, S.con, wc.b0, nc, dump);
    if (l == 0) dump_n[0] = nc;
  }
  float g0[6];
  const float hh = (float)(0.5 * dt);
#pragma unroll
  for (int k = 0; k < 6; ++k) g0[k] = tr ? v[k] + hh * tr[k] : v[k];
  ok = world_solve(M, cfg, wc, lds_ptr(S), nc, q, v, f, dt, g0, qf, vf, wk);
  double err = 0.0;
  if (cfg.mode == 1) {
    float g1[6];
    double qm[7];
    float vm[6];
    const float qh = (float)(0.25 * dt);
#pragma unroll
    for (int k = 0; k < 6; ++k) g1[k] = cfg.warm_start == 2 ? (tr ? v[k] + qh * tr[k] : v[k]) : 0.5f * (v[k] + vf[k]);
    ok = world_solve(M, cfg, wc, lds_ptr(S), nc, q, v, f, 0.5 * dt, g1, qm, vm, wk) && ok;
    int nc2 = world_collide(M, cfg, wc, lds_ptr(S), qm, overflow, wk);
    if (DUMP) {
      dump_contacts(S.seg, S.con, wc.b0, nc2, dump + NA_MAX_CONTACTS * NA_DUMP_FLOATS);
      if (l == 0) dump_n[1] = nc2;
    }
    float g2[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g2[k] = cfg.warm_start ? 2.f * vm[k] - v[k] : vf[k];
    ok = world_solve(M, cfg, wc, lds_ptr(S), nc2, qm, vm, f, 0.5 * dt, g2, qd, vd, wk) && ok;
    double e = 0.0;
    if (own) {
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        double d = fabs(qd[k] - qf[k]);
        e = (d > e || d != d) ? d : e;
      }
    }
    err = wave_max_d(e);
  } else {
    double e = 0.0;
    if (own) {
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        qd[k] = qf[k];
        if (qf[k] != qf[k]) e = qf[k];
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) vd[k] = vf[k];
    }
    err = wave_max_d(e);
  }
  if (isnan(err) || isinf(err)) err = 1.0e10;
  wk[0] += 1;
  return err;
}

// ---------------------------------------------------------------- two-wave tiles
// A tile is one world run by the two waves of a 128-lane workgroup.  The first wave runs the world
// as a single-wave world does; in each attempt the second wave solves the first half step while
// the first solves the full step, both from q_t with the same contacts (the first wave's
// contacts: the second wave copies the first wave's ContactLDS into its own) and separate working
// sets (each wave's own BodyLDS and SolveLDS).  That
// needs the first half step's Newton seed not to depend on the full step's result: warm_start 2.
// Each solve is the same out-of-line world_solve, so a tiled world's results equal its
// single-wave results bit for bit.  The waves meet at two workgroup barriers per attempt; the
// first wave's LDS syncs elsewhere are wave-local.
struct TileHdr {
  int32_t cmd;   // 1: solve the first half step of world w; 0: no more tiles (the second wave leaves)
  int32_t w, nc, ok;
  double h, tol;
  float opt_rel_tol;
  long long wk[NA_WORK];  // the second wave's work counters of its solve
  // handshake of the LDS-flag protocol (workgroups with waves outside the tile): sequence numbers
  // of the last request and the last answer, the first wave's count, and whether the second wave
  // stopped answering (then the first solves the half steps itself, with the same arithmetic)
  int32_t req, done, seq, dead;
};

// The tile's two waves meet at each request and each answer.  In a two-wave workgroup that is a
// workgroup barrier.  Where the workgroup has other waves, which never meet the tile's barriers
// (NA_TILE_FLAGS), the sender publishes a sequence number in LDS after a release fence and the
// receiver polls it, sleeping between reads, then fences (acquire).  Every wait is bounded by the
// constant-rate wall clock (NA_TILE_WAIT_TICKS, 60 s at 100 MHz); a wait that runs out returns
// false and its caller falls back to working alone.
#define NA_TILE_BARRIER 1
#define NA_TILE_FLAGS 2
#define NA_TILE_WAIT_TICKS 6000000000LL
template <int MODE>
__device__ __forceinline__ void tile_signal(int32_t* word, int32_t v) {
  if constexpr (MODE == NA_TILE_FLAGS) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane_id() == 0) __hip_atomic_store(word, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    group_sync();
  }
}
template <int MODE>
__device__ __forceinline__ bool tile_await(int32_t* word, int32_t v) {
  if constexpr (MODE == NA_TILE_FLAGS) {
    const long long t0 = wall_clock64();
    for (int spins = 1; __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != v; ++spins) {
      __builtin_amdgcn_s_sleep(2);
      if ((spins & 255) == 0 && wall_clock64() - t0 > NA_TILE_WAIT_TICKS) return false;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
  } else {
    group_sync();
    return true;
  }
}
// Inputs of the second wave's solve, then its outputs, in the second wave's own segment / solve
// area (read into registers before its solve writes that area; written after the solve).
struct TileIn {
  double q[7][NA_WAVE];
  float v[6][NA_WAVE], f[6][NA_WAVE], g[6][NA_WAVE];
};
struct TileOut {
  double q[7][NA_WAVE];
  float v[6][NA_WAVE];
};
static_assert(sizeof(TileIn) <= sizeof(SolveLDS) && sizeof(TileOut) <= sizeof(SolveLDS),
              "tile handoff fits the solve area");
static_assert(sizeof(ContactLDS) % 16 == 0 && offsetof(WorldLDS, con) % 16 == 0 &&
                  sizeof(MjContactParams<NA_MAX_CONTACTS>) % 16 == 0 && offsetof(WorldLDS, mj) % 16 == 0,
              "contact copies move 16-byte words");
__device__ __forceinline__ TileIn& tile_in(WorldLDS& S) { return *reinterpret_cast<TileIn*>(&S.sol); }
__device__ __forceinline__ TileOut& tile_out(WorldLDS& S) { return *reinterpret_cast<TileOut*>(&S.sol); }

// The first wave's step-doubling attempt of a tiled world (adaptive mode, warm_start 2): the
// semantics and arithmetic of world_attempt, with the first half step solved by the second wave.
template <int MODE>
__device__ __noinline__ double world_attempt_tile(const NaModel& M, const NaConfig& cfg, const WorldCtx& wc,
                                                  lds_ptr_t<WorldLDS> Sp, lds_ptr_t<WorldLDS> S1p, TileHdr& T,
                                                  const double* q, const float* v, const float* f, const float* tr,
                                                  double dt, double* qf, float* vf, double* qd, float* vd, bool& ok,
                                                  int& overflow, int& nc_out, long long* wk) {
  WorldLDS& S = lds_deref<WorldLDS>(Sp);
  WorldLDS& S1 = lds_deref<WorldLDS>(S1p);
  const int l = lane_id();
  const bool own = l < wc.nb;
  int nc = world_collide(M, cfg, wc, lds_ptr(S), q, overflow, wk);
  nc_out = nc;
  float g0[6], g1[6];
  const float hh = (float)(0.5 * dt);
  const float qh = (float)(0.25 * dt);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    g0[k] = tr ? v[k] + hh * tr[k] : v[k];
    g1[k] = tr ? v[k] + qh * tr[k] : v[k];
  }
  double qm[7];
  float vm[6];
  const bool alone = MODE == NA_TILE_FLAGS && T.dead;
  if (!alone) {
    TileIn& in = tile_in(S1);
#pragma unroll
    for (int k = 0; k < 7; ++k) in.q[k][l] = q[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) { in.v[k][l] = v[k]; in.f[k][l] = f[k]; in.g[k][l] = g1[k]; }
    if (l == 0) {
      T.cmd = 1; T.w = wc.w; T.nc = nc; T.h = 0.5 * dt; T.tol = cfg.tol; T.opt_rel_tol = cfg.opt_rel_tol;
    }
  }
  const int32_t seq = T.seq + 1;
  wave_sync();
  if (l == 0) T.seq = seq;
  if (!alone) tile_signal<MODE>(&T.req, seq);  // the second wave starts the first half step
  ok = world_solve(M, cfg, wc, lds_ptr(S), nc, q, v, f, dt, g0, qf, vf, wk);
  if (!alone && tile_await<MODE>(&T.done, seq)) {  // its result is in
    const TileOut& out = tile_out(S1);
#pragma unroll
    for (int k = 0; k < 7; ++k) qm[k] = out.q[k][l];
#pragma unroll
    for (int k = 0; k < 6; ++k) vm[k] = out.v[k][l];
    ok = T.ok && ok;
#pragma unroll
    for (int k = 0; k < NA_WORK; ++k) wk[k] += T.wk[k];
  } else {
    // the second wave did not answer in time (LDS-flag protocol only): the first half step here,
    // the same solve with the same inputs, and no more requests to that wave
    if (l == 0) T.dead = 1;
    ok = world_solve(M, cfg, wc, lds_ptr(S), nc, q, v, f, 0.5 * dt, g1, qm, vm, wk) && ok;
  }
  int nc2 = world_collide(M, cfg, wc, lds_ptr(S), qm, overflow, wk);
  float g2[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) g2[k] = 2.f * vm[k] - v[k];
  ok = world_solve(M, cfg, wc, lds_ptr(S), nc2, qm, vm, f, 0.5 * dt, g2, qd, vd, wk) && ok;
  double e = 0.0;
  if (own) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      double d = fabs(qd[k] - qf[k]);
      e = (d > e || d != d) ? d : e;
    }
  }
  double err = wave_max_d(e);
  if (isnan(err) || isinf(err)) err = 1.0e10;
  wk[0] += 1;
  return err;
}

// The second wave of a tile: solves first half steps on request until told to leave.  Its
// barriers pair with the first wave's: one before each request (or the leave message), one after
// each result.
template <int MODE>
__device__ __noinline__ void tile_helper(const NaModel& M, const NaConfig& cfg, lds_ptr_t<WorldLDS> S0p,
                                         lds_ptr_t<WorldLDS> S1p, TileHdr& T) {
  WorldLDS& S0 = lds_deref<WorldLDS>(S0p);
  WorldLDS& S1 = lds_deref<WorldLDS>(S1p);
  const int l = lane_id();
  for (int32_t seq = 1;; ++seq) {
    if (!tile_await<MODE>(&T.req, seq)) break;
    if (T.cmd != 1) break;
    const WorldCtx wc = world_ctx(M, T.w);
    NaConfig wcfg = cfg;
    wcfg.tol = T.tol;
    wcfg.opt_rel_tol = T.opt_rel_tol;
    const int nc = T.nc;
    const double h = T.h;
    double q[7];
    float v[6], f[6], g[6];
    {
      const TileIn& in = tile_in(S1);
#pragma unroll
      for (int k = 0; k < 7; ++k) q[k] = in.q[k][l];
#pragma unroll
      for (int k = 0; k < 6; ++k) { v[k] = in.v[k][l]; f[k] = in.f[k][l]; g[k] = in.g[k][l]; }
    }
    // the first wave's contacts into this wave's own ContactLDS (and MuJoCo parameters), so the
    // solve reads the same values from this wave's WorldLDS
    {
      constexpr int n4 = (int)(sizeof(ContactLDS) / 16);
      const int4* src = reinterpret_cast<const int4*>(&S0.con);
      int4* dst = reinterpret_cast<int4*>(&S1.con);
      for (int i = l; i < n4; i += NA_WAVE) dst[i] = src[i];
      if (cfg.contact_law == 1) {
        constexpr int m4 = (int)(sizeof(S0.mj) / 16);
        const int4* ms = reinterpret_cast<const int4*>(&S0.mj);
        int4* md = reinterpret_cast<int4*>(&S1.mj);
        for (int i = l; i < m4; i += NA_WAVE) md[i] = ms[i];
      }
    }
    wave_sync();  // the inputs are in registers and the contacts copied before the solve
    long long wk[NA_WORK] = {};
    double qo[7];
    float vo[6];
    const bool ok = world_solve(M, wcfg, wc, lds_ptr(S1), nc, q, v, f, h, g, qo, vo, wk);
    wave_sync();
    TileOut& out = tile_out(S1);
#pragma unroll
    for (int k = 0; k < 7; ++k) out.q[k][l] = qo[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) out.v[k][l] = vo[k];
    if (l == 0) {
      T.ok = ok;
#pragma unroll
      for (int k = 0; k < NA_WORK; ++k) T.wk[k] = wk[k];
    }
    tile_signal<MODE>(&T.done, seq);
  }
}

// Per-world controller scalars held in registers during a frame.
struct WCtrl {
  double dt, dt_half, ideal_dt, sim_time, next_time, last_error, acc_error;
  uint8_t accepted, commit, diverged;
  int substeps, iterations, solve_fail, overflow;
};
__device__ __forceinline__ WCtrl load_ctrl(const Ctrl& K, int w) {
  WCtrl c;
  c.dt = K.dt[w]; c.dt_half = K.dt_half[w]; c.ideal_dt = K.ideal_dt[w]; c.sim_time = K.sim_time[w];
  c.next_time = K.next_time[w]; c.last_error = K.last_error[w]; c.acc_error = K.acc_error[w];
  c.accepted = K.accepted[w]; c.commit = K.commit[w]; c.diverged = K.diverged[w];
  c.substeps = K.substeps[w]; c.iterations = K.iterations[w]; c.solve_fail = K.solve_fail[w];
  c.overflow = K.overflow[w];
  return c;
}
__device__ __forceinline__ void store_ctrl(const Ctrl& K, int w, const WCtrl& c) {
  K.dt[w] = c.dt; K.dt_half[w] = c.dt_half; K.ideal_dt[w] = c.ideal_dt; K.sim_time[w] = c.sim_time;
  K.next_time[w] = c.next_time; K.last_error[w] = c.last_error; K.acc_error[w] = c.acc_error;
  K.accepted[w] = c.accepted; K.commit[w] = c.commit; K.diverged[w] = c.diverged;
  K.substeps[w] = c.substeps; K.iterations[w] = c.iterations; K.solve_fail[w] = c.solve_fail;
  K.overflow[w] = c.overflow;
}

// _open_frame + _seed_dt (SAP) or _apply_dt_cap + _rebase_time + _boundary_advance (MuJoCo)
__device__ __forceinline__ void open_frame(const NaConfig& cfg, WCtrl& c, double dt_outer, double eff_max) {
  if (cfg.controller == 0) {
    double base = c.next_time;
    c.sim_time = c.sim_time - base;
    c.next_time = dt_outer;
    double d = cfg.mode == 0 ? cfg.dt_init : c.ideal_dt;
    d = d < cfg.dt_min ? cfg.dt_min : (d > eff_max ? eff_max : d);
    c.dt = d;
    c.dt_half = 0.5 * d;
    c.substeps = 0;
    c.diverged = 0;
  } else {
    double d = c.ideal_dt;
    d = d < cfg.dt_min ? cfg.dt_min : (d > eff_max ? eff_max : d);
    c.dt = d;
    c.dt_half = 0.5 * d;
    double base = c.next_time;
    c.sim_time = c.sim_time - base;
    c.next_time = c.next_time - base;
    c.next_time = c.next_time + dt_outer;
    c.diverged = 0;
    c.substeps = 0;
  }
  c.iterations = 0;
  c.solve_fail = 0;
  c.overflow = 0;
}

// Controller after an attempt of size `step` with error `err` (ok false: the convex solve
// failed): records the error, advances time and dt, and returns whether to commit the attempt;
// `inc` is set to 1 for an accepted step of positive size.  A failed solve counts as a diverged
// attempt.  The SAP controller also flags it, which ends the frame with the reference's
// RuntimeError (solver_sap_adaptive.py:901-905); the MuJoCo controller has no such error and
// handles it like a NaN state: shrink and retry, at the dt floor hold the state and latch
// `diverged` (solver_mujoco_adaptive.py:120-198).
__device__ __forceinline__ bool controller_update(const NaConfig& cfg, WCtrl& c, double err, bool ok, double step,
                                                  double eff_max, int& inc) {
  if (!ok) {
    err = cfg.divergence_threshold;
    if (cfg.controller == 0) c.solve_fail = 1;
  }
  c.last_error = err;
  bool commit;
  inc = 0;
  if (cfg.controller == 0) {
    commit = adapt_dt<double>(err, c.sim_time, c.next_time, c.dt, c.dt_half, c.ideal_dt, c.diverged, c.acc_error,
                              c.substeps, inc, cfg.mode, cfg.tol, cfg.dt_min, eff_max, cfg.divergence_threshold);
    c.accepted = commit;
  } else {
    calc_adjusted_step<double>(err, c.dt, c.ideal_dt, c.accepted, c.commit, c.diverged, cfg.tol, cfg.dt_min,
                               cfg.divergence_threshold);
    commit = c.commit;
    if (c.accepted) {
      c.sim_time += c.dt;
      c.acc_error = err;
      if (step > 0.0) { c.substeps += 1; inc = 1; }
    }
    if (c.diverged) c.sim_time = c.next_time;
    double d = c.ideal_dt;
    d = d < cfg.dt_min ? cfg.dt_min : (d > eff_max ? eff_max : d);
    c.dt = d;
    c.dt_half = 0.5 * d;
  }
  return commit;
}

// One iteration of the boundary loop for one world (clamp, attempt, controller, commit).
// q/v hold this lane's body state and are updated in place.  trend (warm_start 1 or 2, else null) is
// the per-body acceleration array: read for the full step's Newton seed and, after an accepted
// step, set to (v_new - v_old) / step.  Returns accepted-step count (0/1).
// TILE: 0, or the world is a two-wave tile with that handoff protocol (NA_TILE_BARRIER,
// NA_TILE_FLAGS; S1, T: the second wave's LDS and the tile header).
template <int TILE = 0>
__device__ int world_iteration(const NaModel& M, const NaConfig& cfg, const WorldCtx& wc, WorldLDS& S, WCtrl& c,
                               double* q, float* v, const float* f, float* trend, double eff_max, long long* wk,
                               WorldLDS* S1 = nullptr, TileHdr* T = nullptr) {
  c.iterations += 1;
  const int l = lane_id();
  const bool own = l < wc.nb;
  double rem = c.next_time - c.sim_time;
  if (rem <= 0.0) { c.dt = 0.0; c.dt_half = 0.0; }
  else if (c.dt > rem) { c.dt = rem; c.dt_half = 0.5 * rem; }
  double step = c.dt;
  double qf[7], qd[7];
  float vf[6], vd[6];
  bool ok = true;
  int ovf = 0, nc = 0;
  double err = 0.0;
  if (step > 0.0 && wc.nb > 0) {
    float tr[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) tr[k] = trend && own ? trend[6 * (wc.b0 + l) + k] : 0.f;
    if constexpr (TILE != 0)
      err = world_attempt_tile<TILE>(M, cfg, wc, lds_ptr(S), lds_ptr(*S1), *T, q, v, f, trend ? tr : nullptr, step, qf, vf, qd, vd, ok, ovf,
                               nc, wk);
    else
      err = world_attempt(M, cfg, wc, S, q, v, f, trend ? tr : nullptr, step, qf, vf, qd, vd, ok, ovf, nc, wk);
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) { qf[k] = q[k]; qd[k] = q[k]; }
#pragma unroll
    for (int k = 0; k < 6; ++k) { vf[k] = v[k]; vd[k] = v[k]; }
  }
  if (ovf) c.overflow = 1;
  int inc = 0;
  const bool commit = controller_update(cfg, c, err, ok, step, eff_max, inc);
  if (commit) {
    const double* sq = cfg.mode == 0 ? qf : qd;
    const float* sv = cfg.mode == 0 ? vf : vd;
    if (trend && inc && own) {
      const float inv = (float)(1.0 / step);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const float a = (sv[k] - v[k]) * inv;
        trend[6 * (wc.b0 + l) + k] = isfinite(a) ? a : 0.f;
      }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = sq[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = sv[k];
  }
  return inc;
}


// Load this lane's body (FP64 shadow kept when it still rounds to the caller's f32 pose).
__device__ __forceinline__ void load_body(const WorldCtx& wc, const double* q64, const float* body_q,
                                          const float* body_qd, const float* body_f, double* q, float* v, float* f,
                                          const float* joint_f = nullptr) {
  const int l = lane_id();
  if (l < wc.nb) {
    int gb = wc.b0 + l;
    bool same = true;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      q[k] = q64[7 * gb + k];
      if (body_q) same = same && ((float)q[k] == body_q[7 * gb + k]);
    }
    if (!same) {
#pragma unroll
      for (int k = 0; k < 7; ++k) q[k] = (double)body_q[7 * gb + k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      v[k] = body_qd[6 * gb + k];
      // the body's FREE joint force is a COM wrench in the world frame, like body_f
      f[k] = (body_f ? body_f[6 * gb + k] : 0.f) + (joint_f ? joint_f[6 * gb + k] : 0.f);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = 0.0;
    q[6] = 1.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) { v[k] = 0.f; f[k] = 0.f; }
  }
}
__device__ __forceinline__ void store_body(const WorldCtx& wc, double* q64, float* body_q, float* body_qd,
                                           const double* q, const float* v) {
  const int l = lane_id();
  if (l < wc.nb) {
    int gb = wc.b0 + l;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      q64[7 * gb + k] = q[k];
      body_q[7 * gb + k] = (float)q[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) body_qd[6 * gb + k] = v[k];
  }
}
__device__ __forceinline__ void flush_work(const Ctrl& K, int w, const long long* wk) {
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < NA_WORK; ++k) K.work[(long long)w * NA_WORK + k] += wk[k];
  }
}

// ---------------------------------------------------------------- served boundary call
// One world's boundary frame: iterations from open_frame until the boundary, max_substeps or a
// solve failure.  TILE: the world runs as a two-wave tile (S1, T: the second wave's LDS and the
// tile header).
template <int TILE = 0>
__device__ __forceinline__ void frame_world(const NaModel& M, const NaConfig& cfg, const Ctrl& K, double* q64,
                                            float* body_q, float* body_qd, const float* body_f, const NaControl& U,
                                            double dt_outer, double eff_max, int w, WorldLDS& S,
                                            WorldLDS* S1 = nullptr, TileHdr* T = nullptr) {
  WorldCtx wc = world_ctx(M, w);
  double q[7];
  float v[6], f[6];
  load_body(wc, q64, body_q, body_qd, body_f, q, v, f, U.jf);
  WCtrl c = load_ctrl(K, w);
  open_frame(cfg, c, dt_outer, eff_max);
  long long wk[NA_WORK] = {};
  int acc = 0;
  for (int it = 0; it < cfg.max_substeps; ++it) {
    if (!(c.sim_time < c.next_time)) break;
    acc += world_iteration<TILE>(M, cfg, wc, S, c, q, v, f, K.trend, eff_max, wk, S1, T);
    if (c.solve_fail) break;
  }
  store_body(wc, q64, body_q, body_qd, q, v);
  if (lane_id() == 0) {
    store_ctrl(K, w, c);
    if (acc) atomicAdd(K.cum, (unsigned long long)acc);
  }
  flush_work(K, w, wk);
}

__global__ __launch_bounds__(64) void k_frame(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_q,
                                              float* body_qd, const float* body_f, NaControl U, double dt_outer,
                                              double eff_max, int32_t* queue, const int32_t* order) {
  __shared__ WorldLDS S;
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = atomicAdd(queue, 1);
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= M.W) break;
    const int w = order ? order[slot] : slot;
    frame_world(M, cfg, K, q64, body_q, body_qd, body_f, U, dt_outer, eff_max, w, S);
  }
}

// The boundary call with two-wave tiles: the first n_tiled worlds of the visiting order (the
// previous call's costliest under the cost order) run as tiles, as in k_serve_tile; after them
// both waves of a workgroup take worlds on their own.  Same exit argument as k_serve_tile.
__global__ __launch_bounds__(128) void k_frame_tile(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_q,
                                                    float* body_qd, const float* body_f, NaControl U,
                                                    double dt_outer, double eff_max, int32_t* queue,
                                                    const int32_t* order, int n_tiled) {
  __shared__ WorldLDS S[2];
  __shared__ TileHdr T;
  const int wave = threadIdx.x >> 6;
  if (wave == 1) {
    tile_helper<NA_TILE_BARRIER>(M, cfg, lds_ptr(S[0]), lds_ptr(S[1]), T);
  } else {
    for (;;) {
      int slot = 0;
      if (lane_id() == 0) slot = atomicAdd(queue, 1);
      slot = __shfl(slot, 0, NA_WAVE);
      if (slot >= n_tiled || slot >= M.W) {
        if (lane_id() == 0) T.cmd = 0;
        group_sync();  // the second wave leaves the tile loop
        if (slot < M.W)
          frame_world(M, cfg, K, q64, body_q, body_qd, body_f, U, dt_outer, eff_max, order ? order[slot] : slot,
                      S[0]);
        break;
      }
      frame_world<NA_TILE_BARRIER>(M, cfg, K, q64, body_q, body_qd, body_f, U, dt_outer, eff_max, order ? order[slot] : slot,
                        S[0], &S[1], &T);
    }
  }
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = atomicAdd(queue, 1);
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= M.W) break;
    const int w = order ? order[slot] : slot;
    if (wave == 0)
      frame_world(M, cfg, K, q64, body_q, body_qd, body_f, U, dt_outer, eff_max, w, S[0]);
    else
      frame_world(M, cfg, K, q64, body_q, body_qd, body_f, U, dt_outer, eff_max, w, S[1]);
  }
}

// ---------------------------------------------------------------- rollout serving
// Per-world rollout bookkeeping of the serving loop.  A world runs boundary frames of dt_outer
// (open_frame, iterations until the boundary or max_substeps) until it has completed
// horizon_frames frames or taken horizon_accepted accepted steps (the latter may end it
// inside a frame).  The arithmetic of every iteration is the boundary call's, so a rollout of
// F frames leaves the same state as F boundary calls.
struct ServeSlots {
  const int32_t* horizon_frames;    // [W]
  const int32_t* horizon_accepted;  // [W]
  int32_t* frames_done;             // [W]
  int32_t* accepted;                // [W] accepted steps taken so far
  int32_t* status;                  // [W] 0 running, 1 horizon reached, 2 convex solve failed
  const double* tol;                // [W] per-world error tolerance, or null for cfg.tol
};

// Persistent serving launch over the worlds listed in `active`.  Waves take list entries from an
// atomic counter and advance each world until its horizon, a solve failure, `max_iters`
// iterations (<= 0: no limit) or the time budget (`budget` ticks of the constant-rate wall
// clock, measured from the wave's start); a world stopped early keeps its exact state (FP64 pose
// shadow, controller, frame position) and resumes in a later launch.  With max_iters = 1 every
// launch advances each listed world by one step-doubling iteration: the lock-step form.
// `unfinished` (optional) gets a bit set when any world is still running after its turn.  Every
// wave exits: the loop ends when the list is exhausted or the budget has passed, and each world
// loop runs at most max_substeps iterations per frame.
// One world's turn of a serving launch: boundary frames until the world's horizon, a solve
// failure, max_iters iterations (<= 0: no limit) or the device clock passing t_end.
// TILE: 0, or the world runs as a two-wave tile with that handoff protocol (S1, T: the second
// wave's LDS and the tile header).
template <int TILE = 0>
__device__ __noinline__ void serve_world(const NaModel& M, const NaConfig& cfg, const Ctrl& K, double* q64, float* body_q,
                            float* body_qd, const float* body_f, double dt_outer, double eff_max, const ServeSlots& SV,
                            int w, long long t_end, int max_iters, int32_t* unfinished, lds_ptr_t<WorldLDS> Sp,
                            WorldLDS* S1 = nullptr, TileHdr* T = nullptr) {
  WorldLDS& S = lds_deref<WorldLDS>(Sp);
  WorldCtx wc = world_ctx(M, w);
  NaConfig wcfg = cfg;
  if (SV.tol) {
    wcfg.tol = SV.tol[w];
    wcfg.opt_rel_tol = (float)fmax(1.0e-3 * wcfg.tol, 1.0e-8);
  }
  double q[7];
  float v[6], f[6];
  load_body(wc, q64, body_q, body_qd, body_f, q, v, f);
  WCtrl c = load_ctrl(K, w);
  long long wk[NA_WORK] = {};
  int frames = SV.frames_done[w], accepted = SV.accepted[w], status = 0, acc = 0;
  const int hf = SV.horizon_frames[w], ha = SV.horizon_accepted[w];
  // a world without any finite horizon has nothing to run (keeps every wave's loop finite)
  if (hf >= 0x7fffffff && ha >= 0x7fffffff) status = 1;
  int iters = 0;
  while (status == 0) {
    if (!(c.sim_time < c.next_time) || c.iterations >= wcfg.max_substeps) {
      // at a frame boundary (or the frame's iteration cap): finish, or open the next frame
      if (c.iterations > 0 || c.sim_time > 0.0) ++frames;
      if (frames >= hf) { status = 1; break; }
      open_frame(wcfg, c, dt_outer, eff_max);
    }
    int inc = world_iteration<TILE>(M, wcfg, wc, S, c, q, v, f, K.trend, eff_max, wk, S1, T);
    acc += inc;
    accepted += inc;
    if (c.solve_fail) { status = 2; break; }
    if (accepted >= ha) { status = 1; break; }
    if (++iters == max_iters) {
      // lock-step form: a world whose frame just ended and whose last frame it was is done now
      if (!(c.sim_time < c.next_time) && frames + 1 >= hf) { ++frames; status = 1; }
      break;
    }
    if (wall_clock64() >= t_end) break;
  }
  store_body(wc, q64, body_q, body_qd, q, v);
  if (lane_id() == 0) {
    store_ctrl(K, w, c);
    SV.frames_done[w] = frames;
    SV.accepted[w] = accepted;
    SV.status[w] = status;
    if (acc) atomicAdd(K.cum, (unsigned long long)acc);
    if (status == 0 && unfinished) atomicOr(unfinished, 1);
  }
  flush_work(K, w, wk);
}

__global__ __launch_bounds__(64) void k_serve(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_q,
                                              float* body_qd, const float* body_f, double dt_outer, double eff_max,
                                              int32_t* queue, const int32_t* active, int n_active, ServeSlots SV,
                                              long long budget, int max_iters, int32_t* unfinished) {
  __shared__ WorldLDS S;
  const long long t_end = wall_clock64() + budget;
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = wall_clock64() < t_end ? atomicAdd(queue, 1) : n_active;
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= n_active) break;
    const int w = active[slot];
    if (SV.status[w] != 0) continue;
    serve_world(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished, lds_ptr(S));
  }
}

// Served launch with two-wave tiles: workgroups of two waves.  The first n_tiled entries of the
// list run as tiles (both waves on one world, see world_attempt_tile); once a workgroup's first
// wave takes an entry past them, it tells the second wave to leave the tile loop and both waves
// then take worlds on their own, each with its own WorldLDS, exactly as k_serve's waves do.  The
// list counter only grows, so no tile entry remains once any wave has passed them.  Every wave
// exits: the tile loop ends when the list is exhausted, the budget has passed or an entry past
// the tiles comes up, and each path sends the second wave its leave message once.
__global__ __launch_bounds__(128) void k_serve_tile(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_q,
                                                    float* body_qd, const float* body_f, double dt_outer,
                                                    double eff_max, int32_t* queue, const int32_t* active,
                                                    int n_active, int n_tiled, ServeSlots SV, long long budget,
                                                    int max_iters, int32_t* unfinished) {
  __shared__ WorldLDS S[2];
  __shared__ TileHdr T;
  const int wave = threadIdx.x >> 6;
  const long long t_end = wall_clock64() + budget;
  if (wave == 1) {
    tile_helper<NA_TILE_BARRIER>(M, cfg, lds_ptr(S[0]), lds_ptr(S[1]), T);
  } else {
    for (;;) {
      int slot = 0;
      if (lane_id() == 0) slot = wall_clock64() < t_end ? atomicAdd(queue, 1) : n_active;
      slot = __shfl(slot, 0, NA_WAVE);
      if (slot >= n_tiled || slot >= n_active) {
        if (lane_id() == 0) T.cmd = 0;
        group_sync();  // the second wave leaves the tile loop
        if (slot < n_active) {
          const int w = active[slot];
          if (SV.status[w] == 0)
            serve_world(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters,
                        unfinished, lds_ptr(S[0]));
        }
        break;
      }
      const int w = active[slot];
      if (SV.status[w] != 0) continue;
      serve_world<NA_TILE_BARRIER>(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters,
                        unfinished, lds_ptr(S[0]), &S[1], &T);
    }
  }
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = wall_clock64() < t_end ? atomicAdd(queue, 1) : n_active;
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= n_active) break;
    const int w = active[slot];
    if (SV.status[w] != 0) continue;
    if (wave == 0)
      serve_world(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished,
                  lds_ptr(S[0]));
    else
      serve_world(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished,
                  lds_ptr(S[1]));
  }
}

// ---------------------------------------------------------------- lock-step boundary call
__global__ void k_plain_open(NaModel M, NaConfig cfg, Ctrl K, double dt_outer, double eff_max) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= M.W) return;
  WCtrl c = load_ctrl(K, w);
  open_frame(cfg, c, dt_outer, eff_max);
  store_ctrl(K, w, c);
}
__global__ void k_sync_shadow(int B, double* q64, const float* body_q) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  bool same = true;
  for (int k = 0; k < 7; ++k) same = same && ((float)q64[7 * b + k] == body_q[7 * b + k]);
  if (!same)
    for (int k = 0; k < 7; ++k) q64[7 * b + k] = (double)body_q[7 * b + k];
}
__global__ void k_round_shadow(int B, const double* q64, float* body_q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 7 * B) return;
  body_q[i] = (float)q64[i];
}
__global__ void k_zero_i32(int32_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
// batch_all: a world that has reached its boundary still runs the iteration's two collision
// passes (at its pose, as a zero-length step leaves it) and discards them, as the reference's
// batched CollisionPipeline.collide does for every world (solver_sap_adaptive.py:754,769); its
// state and controller are untouched.
__global__ __launch_bounds__(64) void k_plain_iter(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_qd,
                                                   const float* body_f, NaControl U, double eff_max, int32_t* flag,
                                                   int batch_all) {
  __shared__ WorldLDS S;
  int w = blockIdx.x;
  WCtrl c = load_ctrl(K, w);
  if (!(c.sim_time < c.next_time)) {
    if (batch_all) {
      WorldCtx wc = world_ctx(M, w);
      double q[7];
      float v[6], f[6];
      load_body(wc, q64, nullptr, body_qd, nullptr, q, v, f);
      long long wk[NA_WORK] = {};
      int ovf = 0;
      for (int pass = 0; pass < 2; ++pass) world_collide(M, cfg, wc, lds_ptr(S), q, ovf, wk);
    }
    return;
  }
  WorldCtx wc = world_ctx(M, w);
  const int l = lane_id();
  double q[7];
  float v[6], f[6];
  if (l < wc.nb) {
    int gb = wc.b0 + l;
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = q64[7 * gb + k];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      v[k] = body_qd[6 * gb + k];
      f[k] = (body_f ? body_f[6 * gb + k] : 0.f) + (U.jf ? U.jf[6 * gb + k] : 0.f);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = 0.0;
    q[6] = 1.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) { v[k] = 0.f; f[k] = 0.f; }
  }
  long long wk[NA_WORK] = {};
  int acc = world_iteration(M, cfg, wc, S, c, q, v, f, K.trend, eff_max, wk);
  if (l < wc.nb) {
    int gb = wc.b0 + l;
#pragma unroll
    for (int k = 0; k < 7; ++k) q64[7 * gb + k] = q[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) body_qd[6 * gb + k] = v[k];
  }
  if (l == 0) {
    store_ctrl(K, w, c);
    if (acc) atomicAdd(K.cum, (unsigned long long)acc);
    if (c.solve_fail) atomicOr(flag, 2);
    else if (c.sim_time < c.next_time) atomicOr(flag, 1);
  }
  flush_work(K, w, wk);
}

// ---------------------------------------------------------------- reset
__global__ void k_reset(NaModel M, const uint8_t* mask, double dt_init, Ctrl K) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.W || (mask && !mask[i])) return;
  if (K.trend)  // a reset world has no step history
    for (int j = 6 * M.world_body_start[i]; j < 6 * M.world_body_end[i]; ++j) K.trend[j] = 0.f;
  K.ideal_dt[i] = dt_init;
  K.dt[i] = dt_init;
  K.dt_half[i] = dt_init * 0.5;
  K.sim_time[i] = 0.0;
  K.next_time[i] = 0.0;
  K.diverged[i] = 0;
  K.accepted[i] = 0;
  K.iterations[i] = 0;
}

// ---------------------------------------------------------------- contacts API
#define NA_STAGE_WORDS 20
__global__ __launch_bounds__(64) void k_collide_api(NaModel M, NaConfig cfg, const float* body_q, int per_world,
                                                    int32_t* staging, int32_t* world_counts) {
  __shared__ WorldLDS S;
  int w = blockIdx.x;
  WorldCtx wc = world_ctx(M, w);
  const int l = lane_id();
  double q[7];
  if (l < wc.nb) {
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = body_q[7 * (wc.b0 + l) + k];
  }
  int ovf = 0;
  long long wk[NA_WORK] = {};
  NaConfig c2 = cfg;
  c2.max_contacts = min(per_world, NA_MAX_CONTACTS);
  int nc = world_collide_body(M, c2, wc, S, q, ovf, wk);
  if (l == 0) world_counts[w] = nc;
  // write_contact (collide.py:62-147): body-frame points and offsets
  for (int c = l; c < nc; c += NA_WAVE) {
    int ga = S.seg.gidx[S.con.sa[c]], gb = S.seg.gidx[S.con.sb[c]];
    f3 n = mk(S.con.n[0][c], S.con.n[1][c], S.con.n[2][c]);
    f3 ctr = mk(S.con.p[0][c], S.con.p[1][c], S.con.p[2][c]);
    float d = S.con.dist[c], ra = S.con.ra[c], rb = S.con.rb[c];
    float ma = M.shape_margin[ga], mb = M.shape_margin[gb];
    f3 aw = ctr - n * (0.5f * d + ra);
    f3 bw = ctr + n * (0.5f * d + rb);
    f3 p0 = aw, p1 = bw, o0 = n * (ra + ma), o1 = n * (-(rb + mb));
    int ba = S.con.ba[c], bb = S.con.bb[c];
    if (ba >= 0) {
      m3 R = qmat(S.body.q[3][ba], S.body.q[4][ba], S.body.q[5][ba], S.body.q[6][ba]);
      p0 = mtv(R, aw - mk(S.body.q[0][ba], S.body.q[1][ba], S.body.q[2][ba]));
      o0 = mtv(R, o0);
    }
    if (bb >= 0) {
      m3 R = qmat(S.body.q[3][bb], S.body.q[4][bb], S.body.q[5][bb], S.body.q[6][bb]);
      p1 = mtv(R, bw - mk(S.body.q[0][bb], S.body.q[1][bb], S.body.q[2][bb]));
      o1 = mtv(R, o1);
    }
    int32_t* o = staging + ((long long)w * per_world + c) * NA_STAGE_WORDS;
    o[0] = ga; o[1] = gb;
    float vals[17] = {p0.x, p0.y, p0.z, p1.x, p1.y, p1.z, o0.x, o0.y, o0.z, o1.x, o1.y, o1.z,
                      n.x, n.y, n.z, ra + ma, rb + mb};
#pragma unroll
    for (int k = 0; k < 17; ++k) o[2 + k] = __float_as_int(vals[k]);
  }
}
// exclusive scan of per-world kept counts (single workgroup), total into count[0]
__global__ void k_scan_counts(int W, int per_world, const int32_t* counts, int32_t* offsets, int32_t* total) {
  __shared__ int32_t buf[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < W; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < W ? min(counts[i], per_world) : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < W) offsets[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) total[0] = carry;
}
__global__ void k_compact_contacts(int W, int per_world, const int32_t* staging, const int32_t* counts,
                                   const int32_t* offsets, na_contacts_desc out) {
  int w = blockIdx.x;
  int nc = min(counts[w], per_world);
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    int dst = offsets[w] + c;
    if (dst >= out.capacity) continue;
    const int32_t* s = staging + ((long long)w * per_world + c) * NA_STAGE_WORDS;
    out.shape0[dst] = s[0];
    out.shape1[dst] = s[1];
    const float* f = (const float*)(s + 2);
    for (int k = 0; k < 3; ++k) {
      out.point0[3 * dst + k] = f[k];
      out.point1[3 * dst + k] = f[3 + k];
      out.offset0[3 * dst + k] = f[6 + k];
      out.offset1[3 * dst + k] = f[9 + k];
      out.normal[3 * dst + k] = f[12 + k];
    }
    out.margin0[dst] = f[15];
    out.margin1[dst] = f[16];
  }
}

// Frame-to-frame contact matching (geometry/contact_match.py): one wave per world.  The
// previous frame's contacts of the world are staged in LDS; every new contact looks for previous
// contacts of the same shape pair and takes the closest one whose world midpoint
// 1/2 (world(point0) + world(point1)) moved at most pos_thr and whose normal turned by at most
// acos(dot_thr).  A previous contact goes to one new contact only: claims are packed
// (distance^2 bits << 32 | new contact's index in its world) and resolved with a 64-bit LDS
// atomic min, so the smallest distance wins and ties go to the lower index, independently of
// scheduling.  match: >= 0 index of the previous contact, -1 no previous contact of the pair,
// -2 pair present but no contact within the thresholds or the claim was lost.
#define NA_MATCH_NOT_FOUND (-1)
#define NA_MATCH_BROKEN (-2)
__device__ __forceinline__ f3 contact_world_point(const NaModel& M, const float* body_q, int s, const float* p) {
  int b = M.shape_body[s];
  f3 x = mk(p[0], p[1], p[2]);
  if (b < 0) return x;
  const float* q = body_q + 7 * b;
  return mk(q[0], q[1], q[2]) + mv(qmat(q[3], q[4], q[5], q[6]), x);
}
__global__ __launch_bounds__(64) void k_match_contacts(NaModel M, const float* body_q, na_contacts_desc cur,
                                                       int per_world, const int32_t* cur_off,
                                                       const int32_t* prev_s0, const int32_t* prev_s1,
                                                       const float* prev_mid, const float* prev_n,
                                                       const int32_t* prev_off, const int32_t* prev_cnt,
                                                       int prev_cap, float pos_thr2, float dot_thr, int32_t* match,
                                                       float* cur_mid, int32_t* prev_matched, const float* prev_p0,
                                                       const float* prev_p1, const float* prev_o0,
                                                       const float* prev_o1) {
  __shared__ int ps0[NA_MAX_CONTACTS], ps1[NA_MAX_CONTACTS];
  __shared__ float pm[3][NA_MAX_CONTACTS], pn[3][NA_MAX_CONTACTS];
  __shared__ unsigned long long claim[NA_MAX_CONTACTS];
  const int w = blockIdx.x, l = lane_id();
  const int po = prev_off[w];
  int pc = min(prev_cnt[w], per_world);
  pc = max(0, min(pc, prev_cap - po));
  const int co = cur_off[w];
  int cc = min(cur.world_count_out[w], per_world);
  cc = max(0, min(cc, cur.capacity - co));
  for (int k = l; k < pc; k += NA_WAVE) {
    ps0[k] = prev_s0[po + k];
    ps1[k] = prev_s1[po + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pm[d][k] = prev_mid[3 * (po + k) + d];
      pn[d][k] = prev_n[3 * (po + k) + d];
    }
    claim[k] = ~0ull;
  }
  __syncthreads();
  int best[2] = {-1, -1}, res[2] = {NA_MATCH_NOT_FOUND, NA_MATCH_NOT_FOUND};
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int c = l + g * NA_WAVE;
    if (c >= cc) continue;
    const int i = co + c;
    const int s0 = cur.shape0[i], s1 = cur.shape1[i];
    f3 a = contact_world_point(M, body_q, s0, cur.point0 + 3 * i);
    f3 b = contact_world_point(M, body_q, s1, cur.point1 + 3 * i);
    f3 mid = (a + b) * 0.5f;
    f3 n = mk(cur.normal[3 * i], cur.normal[3 * i + 1], cur.normal[3 * i + 2]);
    cur_mid[3 * i] = mid.x; cur_mid[3 * i + 1] = mid.y; cur_mid[3 * i + 2] = mid.z;
    bool pair = false;
    float bd = 0.f;
    for (int k = 0; k < pc; ++k) {
      if (ps0[k] != s0 || ps1[k] != s1) continue;
      pair = true;
      f3 dm = mid - mk(pm[0][k], pm[1][k], pm[2][k]);
      float d2 = dot3(dm, dm);
      float dn = dot3(n, mk(pn[0][k], pn[1][k], pn[2][k]));
      if (d2 <= pos_thr2 && dn >= dot_thr && (best[g] < 0 || d2 < bd)) { best[g] = k; bd = d2; }
    }
    res[g] = pair ? NA_MATCH_BROKEN : NA_MATCH_NOT_FOUND;
    if (best[g] >= 0)
      atomicMin(&claim[best[g]], ((unsigned long long)__float_as_uint(bd) << 32) | (unsigned long long)c);
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int c = l + g * NA_WAVE;
    if (c >= cc) continue;
    int r = res[g];
    if (best[g] >= 0 && (int)(claim[best[g]] & 0xffffffffull) == c) {
      r = po + best[g];
      prev_matched[po + best[g]] = 1;
    }
    const int i = co + c;
    match[i] = r;
    if (r >= 0 && prev_p0) {
      // sticky: a matched contact keeps the previous frame's body-frame geometry and normal
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        cur.point0[3 * i + d] = prev_p0[3 * r + d];
        cur.point1[3 * i + d] = prev_p1[3 * r + d];
        cur.offset0[3 * i + d] = prev_o0[3 * r + d];
        cur.offset1[3 * i + d] = prev_o1[3 * r + d];
        cur.normal[3 * i + d] = prev_n[3 * r + d];
      }
      const f3 mid = (contact_world_point(M, body_q, cur.shape0[i], cur.point0 + 3 * i) +
                      contact_world_point(M, body_q, cur.shape1[i], cur.point1 + 3 * i)) * 0.5f;
      cur_mid[3 * i] = mid.x; cur_mid[3 * i + 1] = mid.y; cur_mid[3 * i + 2] = mid.z;
    }
  }
}

__global__ __launch_bounds__(64) void k_collide_debug(NaModel M, int broad_phase, const float* body_q, float* lo,
                                                      float* hi, int32_t* pairs, int32_t* pair_count, int max_pairs) {
  __shared__ WorldLDS S;
  int w = blockIdx.x;
  WorldCtx wc = world_ctx(M, w);
  const int l = lane_id();
  if (l < wc.nb) {
#pragma unroll
    for (int k = 0; k < 7; ++k) S.body.q[k][l] = body_q[7 * (wc.b0 + l) + k];
  }
  wave_sync();
  stage_aabb(M, w, wc.nseg, wc.nl, wc.b0, &S.body.q[0][0], NA_MAX_BODIES, S.seg);
  for (int i = l; i < wc.nseg; i += NA_WAVE) {
    if (i >= wc.nl && w != 0) continue;  // global shapes are written once
    int s = S.seg.gidx[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[3 * s + k] = S.seg.lo[k][i];
      hi[3 * s + k] = S.seg.hi[k][i];
    }
  }
  int np = stage_broad(M, wc.w, wc.nseg, wc.nl, broad_phase, S.seg);
  if (l == 0) pair_count[w] = np;
  int m = min(np, min(max_pairs, NA_MAX_PAIRS));
  for (int k = l; k < m; k += NA_WAVE) {
    uint32_t pr = S.seg.pairs[k];
    int a = S.seg.gidx[pr >> 8], b = S.seg.gidx[pr & 0xff];
    pairs[((long long)w * max_pairs + k) * 2] = min(a, b);
    pairs[((long long)w * max_pairs + k) * 2 + 1] = max(a, b);
  }
}

// ---------------------------------------------------------------- parity entry kernels
__global__ __launch_bounds__(64) void k_substep(NaModel M, NaConfig cfg, const double* q64, const float* qd,
                                                const float* f, const double* h, const double* guess, double* q_out,
                                                float* qd_out, double* v_solution, int32_t* stats) {
  __shared__ WorldLDS S;
  int w = blockIdx.x;
  WorldCtx wc = world_ctx(M, w);
  const int l = lane_id();
  double q[7], qo[7];
  float v[6], ff[6], vo[6];
  load_body(wc, q64, nullptr, qd, f, q, v, ff);
  if (l < wc.nb) {
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = q64[7 * (wc.b0 + l) + k];
  }
  if (wc.nb == 0) return;
  long long wk[NA_WORK] = {};
  int ovf = 0;
  int nc = world_collide(M, cfg, wc, lds_ptr(S), q, ovf, wk);
  SolveStats st;
  double g64[6], v64[6];
  const bool own = l < wc.nb;
  if (guess) {
#pragma unroll
    for (int k = 0; k < 6; ++k) g64[k] = own ? guess[6 * (wc.b0 + l) + k] : 0.0;
  }
  bool ok = world_solve(M, cfg, wc, lds_ptr(S), nc, q, v, ff, h[w], v, qo, vo, wk, &st,
                        guess ? g64 : nullptr, v64);
  if (own) {
    int gb = wc.b0 + l;
#pragma unroll
    for (int k = 0; k < 7; ++k) q_out[7 * gb + k] = qo[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) qd_out[6 * gb + k] = vo[k];
    if (v_solution) {
#pragma unroll
      for (int k = 0; k < 6; ++k) v_solution[6 * gb + k] = v64[k];
    }
  }
  if (l == 0) {
    stats[6 * w + 0] = st.iterations;
    stats[6 * w + 1] = st.cg_iterations;
    stats[6 * w + 2] = ok ? 1 : 0;
    stats[6 * w + 3] = nc;
    stats[6 * w + 4] = st.reason;
    stats[6 * w + 5] = __float_as_int(st.residual);
  }
}
__global__ __launch_bounds__(64) void k_attempt(NaModel M, NaConfig cfg, const double* q64, const float* qd,
                                                const float* f, const double* dt, double* q_full, float* qd_full,
                                                double* q_dbl, float* qd_dbl, double* err, int32_t* solve_ok,
                                                int32_t* n_contacts, float* cdump, int32_t* cdump_n,
                                                const float* trend) {
  __shared__ WorldLDS S;
  int w = blockIdx.x;
  WorldCtx wc = world_ctx(M, w);
  const int l = lane_id();
  double q[7], qf[7], qd2[7];
  float v[6], ff[6], vf[6], vd[6];
  load_body(wc, q64, nullptr, qd, f, q, v, ff);
  if (l < wc.nb) {
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = q64[7 * (wc.b0 + l) + k];
  }
  if (wc.nb == 0) {
    if (l == 0) {
      err[w] = 0.0; solve_ok[w] = 1; n_contacts[w] = 0;
      if (cdump_n) { cdump_n[2 * w] = 0; cdump_n[2 * w + 1] = 0; }
    }
    return;
  }
  long long wk[NA_WORK] = {};
  bool ok = true;
  int ovf = 0, nc = 0;
  double e;
  float tr[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) tr[k] = trend && l < wc.nb ? trend[6 * (wc.b0 + l) + k] : 0.f;
  const float* trp = trend ? tr : nullptr;
  if (cdump) {
    if (l == 0) cdump_n[2 * w + 1] = 0;  // fixed mode has no midpoint pass
    e = world_attempt<true>(M, cfg, wc, S, q, v, ff, trp, dt[w], qf, vf, qd2, vd, ok, ovf, nc, wk,
                            cdump + (long long)w * 2 * NA_MAX_CONTACTS * NA_DUMP_FLOATS, cdump_n + 2 * w);
  } else {
    e = world_attempt(M, cfg, wc, S, q, v, ff, trp, dt[w], qf, vf, qd2, vd, ok, ovf, nc, wk);
  }
  if (l < wc.nb) {
    int gb = wc.b0 + l;
#pragma unroll
    for (int k = 0; k < 7; ++k) { q_full[7 * gb + k] = qf[k]; q_dbl[7 * gb + k] = qd2[k]; }
#pragma unroll
    for (int k = 0; k < 6; ++k) { qd_full[6 * gb + k] = vf[k]; qd_dbl[6 * gb + k] = vd[k]; }
  }
  if (l == 0) {
    err[w] = e;
    solve_ok[w] = ok ? 1 : 0;
    n_contacts[w] = nc;
  }
}

// ---------------------------------------------------------------- articulated worlds
// One iteration of the boundary loop for an articulated world (clamp, attempt, controller,
// commit); q / v are this lane's coordinate and dof velocity (FP64), updated in place.
__device__ int art_iteration(const NaModel& M, const NaConfig& cfg, const ArtWorld& aw, ArtLDS& A, WCtrl& c,
                             double& q, double& v, double eff_max, long long* wk) {
  c.iterations += 1;
  double rem = c.next_time - c.sim_time;
  if (rem <= 0.0) { c.dt = 0.0; c.dt_half = 0.0; }
  else if (c.dt > rem) { c.dt = rem; c.dt_half = 0.5 * rem; }
  const double step = c.dt;
  double qf = q, vf = v, qd = q, vd = v;
  bool ok = true;
  int ovf = 0, nc = 0;
  double err = 0.0;
  if (step > 0.0 && aw.nb > 0) err = art_attempt(M, cfg, aw, lds_ptr(A), q, v, step, qf, vf, qd, vd, ok, ovf, nc, wk);
  if (ovf) c.overflow = 1;
  int inc = 0;
  if (controller_update(cfg, c, err, ok, step, eff_max, inc)) {
    q = cfg.mode == 0 ? qf : qd;
    v = cfg.mode == 0 ? vf : vd;
  }
  return inc;
}

// Body poses and COM / angular velocities of the world at coordinates q, velocities v (eval_fk).
__device__ void art_store_bodies(const NaModel& M, const ArtWorld& aw, ArtLDS& A, double q, double v, float* body_q,
                                 float* body_qd) {
  const int l = lane_id();
  art_kinematics(M, aw, lds_ptr(A), q);
  if (l < aw.nv) A.x[l] = v;
  wave_sync();
  art_body_velocity(aw, A);
  if (l < aw.nb) {
    const int gb = aw.b0 + l;
#pragma unroll
    for (int k = 0; k < 7; ++k) body_q[7 * gb + k] = A.bqf[k][l];
    if (body_qd) {
      double w[3] = {A.V[l][0], A.V[l][1], A.V[l][2]}, c[3] = {A.cb[l][0], A.cb[l][1], A.cb[l][2]}, wc[3];
      dcross3(w, c, wc);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        body_qd[6 * gb + k] = (float)(A.V[l][3 + k] + wc[k]);
        body_qd[6 * gb + 3 + k] = (float)w[k];
      }
    }
  }
}

// Served boundary call of articulated worlds: persistent waves over the world queue, each world
// marched to its boundary on the device with FP64 coordinates (shadow q64 kept while it rounds to
// the caller's joint_q) and velocities.
__global__ __launch_bounds__(64) void k_art_frame(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* joint_q,
                                                  float* joint_qd, float* body_q, float* body_qd, const float* body_f,
                                                  NaControl U, double dt_outer, double eff_max, int32_t* queue,
                                                  const int32_t* order) {
  __shared__ ArtLDS A;
  const int l = lane_id();
  for (;;) {
    int slot = 0;
    if (l == 0) slot = atomicAdd(queue, 1);
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= M.W) break;
    const int w = order ? order[slot] : slot;
    ArtWorld aw = art_world(M, w);
    if (aw.nb == 0) continue;
    art_load(M, aw, A, U, body_f);
    double q = 0.0, v = 0.0;
    int same = 1;
    if (l < aw.nq) {
      q = q64[aw.q0 + l];
      same = (float)q == joint_q[aw.q0 + l] ? 1 : 0;
    }
    same = wave_min_i(same);
    if (!same && l < aw.nq) q = (double)joint_q[aw.q0 + l];
    if (l < aw.nv) v = (double)joint_qd[aw.d0 + l];
    WCtrl c = load_ctrl(K, w);
    open_frame(cfg, c, dt_outer, eff_max);
    long long wk[NA_WORK] = {};
    int acc = 0;
    for (int it = 0; it < cfg.max_substeps; ++it) {
      if (!(c.sim_time < c.next_time)) break;
      acc += art_iteration(M, cfg, aw, A, c, q, v, eff_max, wk);
      if (c.solve_fail) break;
    }
    if (l < aw.nq) {
      q64[aw.q0 + l] = q;
      joint_q[aw.q0 + l] = (float)q;
    }
    if (l < aw.nv) joint_qd[aw.d0 + l] = (float)v;
    if (body_q) art_store_bodies(M, aw, A, q, v, body_q, body_qd);
    if (l == 0) {
      store_ctrl(K, w, c);
      if (acc) atomicAdd(K.cum, (unsigned long long)acc);
    }
    flush_work(K, w, wk);
  }
}

// Persistent serving launch over articulated worlds (the articulated counterpart of k_serve, same
// horizons, budget and lock-step semantics).  Coordinates and velocities resume from FP64 shadows
// (q64 per coordinate, qd64 per dof) while they still round to the caller's joint_q / joint_qd, so a
// world interrupted by the budget continues exactly; at every frame boundary the velocity is
// rounded to float32 as a boundary call's joint_qd would be, so a rollout of F frames leaves the
// same state as F calls of na_step_articulated.
// Caller-owned state of articulated worlds in a serving launch.
struct ArtServeState {
  double* q64;      // FP64 coordinate shadow
  double* qd64;     // FP64 velocity shadow
  float* joint_q;
  float* joint_qd;
  float* body_q;    // forward kinematics output (may be null)
  float* body_qd;
  const float* body_f;  // applied COM wrenches (may be null)
};

// One articulated world's turn of a serving launch (the semantics of serve_world).
__device__ __noinline__ void serve_world_art(const NaModel& M, const NaConfig& cfg, const Ctrl& K, const ArtServeState& X,
                                double dt_outer, double eff_max, const ServeSlots& SV, int w, long long t_end,
                                int max_iters, int32_t* unfinished, lds_ptr_t<ArtLDS> Ap) {
  ArtLDS& A = lds_deref<ArtLDS>(Ap);
  const int l = lane_id();
  ArtWorld aw = art_world(M, w);
  if (aw.nb == 0) return;
  art_load(M, aw, A, NaControl{}, X.body_f);
  NaConfig wcfg = cfg;
  if (SV.tol) {
    wcfg.tol = SV.tol[w];
    wcfg.opt_rel_tol = (float)fmax(1.0e-3 * wcfg.tol, 1.0e-8);
  }
  double q = 0.0, v = 0.0;
  int same = 1;
  if (l < aw.nq) {
    q = X.q64[aw.q0 + l];
    same = (float)q == X.joint_q[aw.q0 + l] ? 1 : 0;
  }
  if (l < aw.nv) {
    v = X.qd64[aw.d0 + l];
    same = same && (float)v == X.joint_qd[aw.d0 + l];
  }
  same = wave_min_i(same ? 1 : 0);
  if (!same) {
    q = l < aw.nq ? (double)X.joint_q[aw.q0 + l] : 0.0;
    v = l < aw.nv ? (double)X.joint_qd[aw.d0 + l] : 0.0;
  }
  WCtrl c = load_ctrl(K, w);
  long long wk[NA_WORK] = {};
  int frames = SV.frames_done[w], accepted = SV.accepted[w], status = 0, acc = 0;
  const int hf = SV.horizon_frames[w], ha = SV.horizon_accepted[w];
  if (hf >= 0x7fffffff && ha >= 0x7fffffff) status = 1;
  int iters = 0;
  while (status == 0) {
    if (!(c.sim_time < c.next_time) || c.iterations >= wcfg.max_substeps) {
      if (c.iterations > 0 || c.sim_time > 0.0) ++frames;
      if (frames >= hf) { status = 1; break; }
      open_frame(wcfg, c, dt_outer, eff_max);
      v = (double)(float)v;  // the boundary state's velocity is float32
    }
    const int inc = art_iteration(M, wcfg, aw, A, c, q, v, eff_max, wk);
    acc += inc;
    accepted += inc;
    if (c.solve_fail) { status = 2; break; }
    if (accepted >= ha) { status = 1; break; }
    if (++iters == max_iters) {
      if (!(c.sim_time < c.next_time) && frames + 1 >= hf) { ++frames; status = 1; }
      break;
    }
    if (wall_clock64() >= t_end) break;
  }
  if (l < aw.nq) {
    X.q64[aw.q0 + l] = q;
    X.joint_q[aw.q0 + l] = (float)q;
  }
  if (l < aw.nv) {
    X.qd64[aw.d0 + l] = v;
    X.joint_qd[aw.d0 + l] = (float)v;
  }
  if (X.body_q) art_store_bodies(M, aw, A, q, v, X.body_q, X.body_qd);
  if (l == 0) {
    store_ctrl(K, w, c);
    SV.frames_done[w] = frames;
    SV.accepted[w] = accepted;
    SV.status[w] = status;
    if (acc) atomicAdd(K.cum, (unsigned long long)acc);
    if (status == 0 && unfinished) atomicOr(unfinished, 1);
  }
  flush_work(K, w, wk);
}

// Persistent serving launch over articulated worlds (the articulated counterpart of k_serve, same
// horizons, budget and lock-step semantics).  Coordinates and velocities resume from FP64 shadows
// (q64 per coordinate, qd64 per dof) while they still round to the caller's joint_q / joint_qd, so a
// world interrupted by the budget continues exactly; at every frame boundary the velocity is
// rounded to float32 as a boundary call's joint_qd would be, so a rollout of F frames leaves the
// same state as F calls of na_step_articulated.
__global__ __launch_bounds__(64) void k_art_serve(NaModel M, NaConfig cfg, Ctrl K, ArtServeState X, double dt_outer,
                                                  double eff_max, int32_t* queue, const int32_t* active, int n_active,
                                                  ServeSlots SV, long long budget, int max_iters,
                                                  int32_t* unfinished) {
  __shared__ ArtLDS A;
  const long long t_end = wall_clock64() + budget;
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = wall_clock64() < t_end ? atomicAdd(queue, 1) : n_active;
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= n_active) break;
    const int w = active[slot];
    if (SV.status[w] != 0) continue;
    serve_world_art(M, cfg, K, X, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished, lds_ptr(A));
  }
}

// Serving launch over a pool that holds free-body and articulated worlds: every wave takes the
// next listed world of either kind from one queue, so the two kinds share the device and every
// world advances continuously.  A world is articulated when its first body carries a joint
// (joint_type != 0).  Each world's arithmetic is that of k_serve / k_art_serve.
__global__ __launch_bounds__(64) void k_serve_mixed(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_q,
                                                    float* body_qd, const float* body_f, ArtServeState X,
                                                    double dt_outer, double eff_max, int32_t* queue,
                                                    const int32_t* active, int n_active, ServeSlots SV,
                                                    long long budget, int max_iters, int32_t* unfinished) {
  // one LDS block, viewed as the free-body or the articulated working set
  constexpr size_t kMixBytes = sizeof(WorldLDS) > sizeof(ArtLDS) ? sizeof(WorldLDS) : sizeof(ArtLDS);
  __shared__ __align__(16) char mix_lds[kMixBytes];
  WorldLDS& Uf = *reinterpret_cast<WorldLDS*>(mix_lds);
  ArtLDS& Ua = *reinterpret_cast<ArtLDS*>(mix_lds);
  const long long t_end = wall_clock64() + budget;
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = wall_clock64() < t_end ? atomicAdd(queue, 1) : n_active;
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= n_active) break;
    const int w = active[slot];
    if (SV.status[w] != 0) continue;
    const int b0 = M.world_body_start[w];
    const bool art = M.world_body_end[w] > b0 && M.joint_type[b0] != 0;
    if (art) serve_world_art(M, cfg, K, X, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished, lds_ptr(Ua));
    else serve_world(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished, lds_ptr(Uf));
  }
}

// Mixed serving pool with two-wave tiles: workgroups of three waves, one LDS block each (as in
// k_serve_mixed, so still three waves per CU).  The list's first n_tiled entries (free-body worlds
// with the most work left) are taken, through their own counter queue[0], only by the first wave
// of each workgroup, which runs each as a tile with the second wave (see world_attempt_tile); the
// third wave takes the rest of the list through queue[1] from the start, and the first two join it
// once the tile entries are exhausted.  An articulated entry among the tiles runs on the first wave
// alone.  The tile pair meets through the LDS-flag handshake (NA_TILE_FLAGS), since the third wave
// never meets their barriers.  Every wave exits: each loop ends when its counter passes its list or
// the budget has passed, the first wave sends the second its leave message once, on leaving the tile
// loop, and every wait of the handshake is bounded.
__device__ __forceinline__ void serve_mixed_world(const NaModel& M, const NaConfig& cfg, const Ctrl& K, double* q64,
                                                  float* body_q, float* body_qd, const float* body_f,
                                                  const ArtServeState& X, double dt_outer, double eff_max,
                                                  const ServeSlots& SV, int w, long long t_end, int max_iters,
                                                  int32_t* unfinished, char* block) {
  const int b0 = M.world_body_start[w];
  const bool art = M.world_body_end[w] > b0 && M.joint_type[b0] != 0;
  if (art)
    serve_world_art(M, cfg, K, X, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished,
                    lds_ptr(*reinterpret_cast<ArtLDS*>(block)));
  else
    serve_world(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished,
                lds_ptr(*reinterpret_cast<WorldLDS*>(block)));
}

__global__ __launch_bounds__(192) void k_serve_mixed_tile(NaModel M, NaConfig cfg, Ctrl K, double* q64, float* body_q,
                                                         float* body_qd, const float* body_f, ArtServeState X,
                                                         double dt_outer, double eff_max, int32_t* queue,
                                                         const int32_t* active, int n_active, int n_tiled,
                                                         ServeSlots SV, long long budget, int max_iters,
                                                         int32_t* unfinished) {
  constexpr size_t kMixBytes = sizeof(WorldLDS) > sizeof(ArtLDS) ? sizeof(WorldLDS) : sizeof(ArtLDS);
  __shared__ __align__(16) char mix0[kMixBytes];
  __shared__ __align__(16) char mix1[kMixBytes];
  __shared__ __align__(16) char mix2[kMixBytes];
  __shared__ TileHdr T;
  WorldLDS& W0 = *reinterpret_cast<WorldLDS*>(mix0);
  WorldLDS& W1 = *reinterpret_cast<WorldLDS*>(mix1);
  const int wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) { T.req = 0; T.done = 0; T.seq = 0; T.dead = 0; }
  group_sync();  // the only workgroup barrier: every wave passes it once, before the roles split
  const long long t_end = wall_clock64() + budget;
  const int n_tiles = n_tiled < n_active ? n_tiled : n_active;
  if (wave == 1) {
    tile_helper<NA_TILE_FLAGS>(M, cfg, lds_ptr(W0), lds_ptr(W1), T);
  } else if (wave == 0) {
    for (;;) {
      int slot = 0;
      if (lane_id() == 0) slot = wall_clock64() < t_end ? atomicAdd(&queue[0], 1) : n_tiles;
      slot = __shfl(slot, 0, NA_WAVE);
      if (slot >= n_tiles) {
        if (lane_id() == 0) T.cmd = 0;
        const int32_t seq = T.seq + 1;
        wave_sync();
        tile_signal<NA_TILE_FLAGS>(&T.req, seq);  // the second wave leaves the tile loop
        break;
      }
      const int w = active[slot];
      if (SV.status[w] != 0) continue;
      const int b0 = M.world_body_start[w];
      if (M.world_body_end[w] > b0 && M.joint_type[b0] != 0)
        serve_world_art(M, cfg, K, X, dt_outer, eff_max, SV, w, t_end, max_iters, unfinished,
                        lds_ptr(*reinterpret_cast<ArtLDS*>(mix0)));
      else
        serve_world<NA_TILE_FLAGS>(M, cfg, K, q64, body_q, body_qd, body_f, dt_outer, eff_max, SV, w, t_end, max_iters,
                          unfinished, lds_ptr(W0), &W1, &T);
    }
  }
  for (;;) {
    int slot = 0;
    if (lane_id() == 0) slot = wall_clock64() < t_end ? n_tiles + atomicAdd(&queue[1], 1) : n_active;
    slot = __shfl(slot, 0, NA_WAVE);
    if (slot >= n_active) break;
    const int w = active[slot];
    if (SV.status[w] != 0) continue;
    if (wave == 0)
      serve_mixed_world(M, cfg, K, q64, body_q, body_qd, body_f, X, dt_outer, eff_max, SV, w, t_end, max_iters,
                        unfinished, mix0);
    else if (wave == 1)
      serve_mixed_world(M, cfg, K, q64, body_q, body_qd, body_f, X, dt_outer, eff_max, SV, w, t_end, max_iters,
                        unfinished, mix1);
    else
      serve_mixed_world(M, cfg, K, q64, body_q, body_qd, body_f, X, dt_outer, eff_max, SV, w, t_end, max_iters,
                        unfinished, mix2);
  }
}

// Parity entry kernels: one inner step / one step-doubling attempt per articulated world.
__global__ __launch_bounds__(64) void k_art_substep(NaModel M, NaConfig cfg, const double* q64, const double* qd64,
                                                    const double* h, const float* body_f, NaControl U, double* q_out,
                                                    double* qd_out, int32_t* stats) {
  __shared__ ArtLDS A;
  const int w = blockIdx.x;
  const int l = lane_id();
  ArtWorld aw = art_world(M, w);
  if (aw.nb == 0) {
    if (l < 4) stats[4 * w + l] = l == 2 ? 1 : 0;
    return;
  }
  art_load(M, aw, A, U, body_f);
  const double q = l < aw.nq ? q64[aw.q0 + l] : 0.0;
  const double v = l < aw.nv ? qd64[aw.d0 + l] : 0.0;
  long long wk[NA_WORK] = {};
  int ovf = 0;
  double mdiag;
  art_kinematics(M, aw, lds_ptr(A), q);
  const double cb = art_dynamics(M, aw, A, v, mdiag);
  const int nc = art_collide(M, cfg, aw, lds_ptr(A), ovf, wk);
  ArtStats st;
  double vo, qo;
  const bool ok = art_step(cfg, aw, lds_ptr(A), nc, h[w], v, cb, mdiag, v, vo, qo, st, wk);
  if (l < aw.nq) q_out[aw.q0 + l] = qo;
  if (l < aw.nv) qd_out[aw.d0 + l] = vo;
  if (l == 0) {
    stats[4 * w + 0] = st.iterations;
    stats[4 * w + 1] = st.factorizations;
    stats[4 * w + 2] = ok ? 1 : 0;
    stats[4 * w + 3] = nc;
  }
}
__global__ __launch_bounds__(64) void k_art_attempt(NaModel M, NaConfig cfg, const double* q64, const double* qd64,
                                                    const double* dt, const float* body_f, NaControl U,
                                                    double* q_full, double* qd_full, double* q_dbl, double* qd_dbl,
                                                    double* err, int32_t* solve_ok, int32_t* n_contacts,
                                                    float* cdump, int32_t* cdump_n) {
  __shared__ ArtLDS A;
  const int w = blockIdx.x;
  const int l = lane_id();
  ArtWorld aw = art_world(M, w);
  if (aw.nb == 0) {
    if (l == 0) {
      err[w] = 0.0; solve_ok[w] = 1; n_contacts[w] = 0;
      if (cdump_n) { cdump_n[2 * w] = 0; cdump_n[2 * w + 1] = 0; }
    }
    return;
  }
  art_load(M, aw, A, U, body_f);
  const double q = l < aw.nq ? q64[aw.q0 + l] : 0.0;
  const double v = l < aw.nv ? qd64[aw.d0 + l] : 0.0;
  long long wk[NA_WORK] = {};
  bool ok = true;
  int ovf = 0, nc = 0;
  double qf, vf, qd, vd;
  double e;
  if (cdump) {
    if (l == 0) cdump_n[2 * w + 1] = 0;  // fixed mode has no midpoint pass
    e = art_attempt<true>(M, cfg, aw, lds_ptr(A), q, v, dt[w], qf, vf, qd, vd, ok, ovf, nc, wk,
                          cdump + (long long)w * 2 * NA_MAX_CONTACTS * NA_DUMP_FLOATS, cdump_n + 2 * w);
  } else {
    e = art_attempt(M, cfg, aw, lds_ptr(A), q, v, dt[w], qf, vf, qd, vd, ok, ovf, nc, wk);
  }
  if (l < aw.nq) { q_full[aw.q0 + l] = qf; q_dbl[aw.q0 + l] = qd; }
  if (l < aw.nv) { qd_full[aw.d0 + l] = vf; qd_dbl[aw.d0 + l] = vd; }
  if (l == 0) {
    err[w] = e;
    solve_ok[w] = ok ? 1 : 0;
    n_contacts[w] = nc;
  }
}

// ---------------------------------------------------------------- host side
static inline hipStream_t S_(void* s) { return (hipStream_t)s; }
static inline int blocks(long long n, int t) { return (int)((n + t - 1) / t); }

extern "C" int na_device_info(char* name, int name_len, int* cu_count, int* lds_per_block) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    strncpy(name, p.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (lds_per_block) *lds_per_block = (int)sizeof(WorldLDS);
  return 0;
}

// Articulation rules of the wave-per-world kernels: one joint per body in body order, revolute or
// FREE (root only), parents first and inside the world, at most NA_ART_NB bodies, NA_ART_NQ
// coordinates and NA_ART_ND dofs per world.  Returns an error message or "".
static std::string validate_articulation(const NaModel& M, const std::vector<int>& wbs, const std::vector<int>& wbe) {
  if (!M.joint_parent || !M.joint_q_start || !M.joint_qd_start || !M.joint_X_p || !M.joint_X_c || !M.joint_axis ||
      !M.joint_target_pos || !M.joint_target_ke || !M.joint_target_kd || !M.joint_armature)
    return "na_model_create: joint_type given without every other joint array";
  const int B = M.B;
  std::vector<int> jt(B > 0 ? B : 1), jp(B > 0 ? B : 1), qs(B + 1), ds(B + 1);
  if (B > 0 && (hipMemcpy(jt.data(), M.joint_type, sizeof(int) * B, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(jp.data(), M.joint_parent, sizeof(int) * B, hipMemcpyDeviceToHost) != hipSuccess))
    return "na_model_create: could not read the joint arrays (device pointers?)";
  if (hipMemcpy(qs.data(), M.joint_q_start, sizeof(int) * (B + 1), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(ds.data(), M.joint_qd_start, sizeof(int) * (B + 1), hipMemcpyDeviceToHost) != hipSuccess)
    return "na_model_create: could not read joint_q_start / joint_qd_start";
  for (int w = 0; w < M.W; ++w) {
    const int b0 = wbs[w], b1 = wbe[w];
    if (b1 - b0 > NA_ART_NB) return "na_model_create: an articulated world has more than 16 bodies";
    if (qs[b1] - qs[b0] > NA_ART_NQ) return "na_model_create: an articulated world has more than 32 coordinates";
    if (ds[b1] - ds[b0] > NA_ART_ND) return "na_model_create: an articulated world has more than 32 dofs";
    for (int b = b0; b < b1; ++b) {
      const int t = jt[b], p = jp[b];
      const int nq = qs[b + 1] - qs[b], nd = ds[b + 1] - ds[b];
      if (t == NA_JT_FREE) {
        if (p != -1) return "na_model_create: a FREE joint must have the world as parent";
        if (nq != 7 || nd != 6) return "na_model_create: a FREE joint has 7 coordinates and 6 dofs";
      } else if (t == NA_JT_REVOLUTE) {
        if (nq != 1 || nd != 1) return "na_model_create: a revolute joint has 1 coordinate and 1 dof";
        if (p != -1 && (p < b0 || p >= b)) return "na_model_create: a joint parent must precede its child in its world";
      } else {
        return "na_model_create: joint types must be 1 (revolute) or 4 (FREE)";
      }
    }
  }
  return std::string();
}

extern "C" int na_model_create(const na_model_desc* d, na_model** out) {
  if (!d || !out) return fail("na_model_create: null argument");
  if (d->world_count <= 0) return fail("na_model_create: world_count must be > 0");
  na_model* m = new na_model();
  NaModel& M = m->M;
  M.W = d->world_count; M.B = d->body_count; M.S = d->shape_count; M.G = d->global_count; M.F = d->filter_count;
  M.world_body_start = d->world_body_start; M.world_shape_start = d->world_shape_start;
  M.world_body_end = d->world_body_end ? d->world_body_end : d->world_body_start + 1;
  M.world_shape_end = d->world_shape_end ? d->world_shape_end : d->world_shape_start + 1;
  M.global_shapes = d->global_shapes;
  M.body_com = d->body_com; M.body_mass = d->body_mass; M.body_inv_mass = d->body_inv_mass;
  M.body_inertia = d->body_inertia; M.body_inv_inertia = d->body_inv_inertia;
  M.shape_type = d->shape_type; M.shape_body = d->shape_body; M.shape_group = d->shape_collision_group;
  M.shape_world = d->shape_world; M.shape_xform = d->shape_transform; M.shape_scale = d->shape_scale;
  M.shape_margin = d->shape_margin; M.shape_gap = d->shape_gap; M.shape_ke = d->shape_material_ke;
  M.shape_mu = d->shape_material_mu; M.shape_radius = d->shape_collision_radius; M.filter_pairs = d->filter_pairs;
  M.gx = d->gravity[0]; M.gy = d->gravity[1]; M.gz = d->gravity[2];
  M.P = d->explicit_pair_count; M.pair_list = d->explicit_pairs; M.world_pair_start = d->world_pair_start;
  if (M.P < 0 || (M.P > 0 && (!M.pair_list || !M.world_pair_start))) {
    delete m;
    return fail("na_model_create: explicit pairs need explicit_pairs [P,2] and world_pair_start [W+1]");
  }
  // validate per-world sizes against the wave-per-world limits
  std::string err;
  // the end arrays are either separate [W] arrays or the start arrays shifted by one
  std::vector<int> wbs(M.W + 1), wss(M.W + 1), wbe(M.W + 1), wse(M.W + 1);
  hipError_t e1 = hipMemcpy(wbs.data(), d->world_body_start, sizeof(int) * M.W, hipMemcpyDeviceToHost);
  hipError_t e2 = hipMemcpy(wss.data(), d->world_shape_start, sizeof(int) * M.W, hipMemcpyDeviceToHost);
  hipError_t e3 = hipMemcpy(wbe.data(), M.world_body_end, sizeof(int) * M.W, hipMemcpyDeviceToHost);
  hipError_t e4 = hipMemcpy(wse.data(), M.world_shape_end, sizeof(int) * M.W, hipMemcpyDeviceToHost);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
    delete m;
    return fail("na_model_create: could not read the world body/shape ranges (device pointers?)");
  }
  m->max_nb = 0; m->max_nl = 0;
  for (int w = 0; w < M.W; ++w) {
    if (wbe[w] < wbs[w] || wse[w] < wss[w]) { delete m; return fail("na_model_create: a world range ends before it starts"); }
    m->max_nb = wbe[w] - wbs[w] > m->max_nb ? wbe[w] - wbs[w] : m->max_nb;
    m->max_nl = wse[w] - wss[w] > m->max_nl ? wse[w] - wss[w] : m->max_nl;
  }
  m->max_seg = m->max_nl + M.G;
  m->articulated = d->joint_type != nullptr;
  M.joint_type = d->joint_type; M.joint_parent = d->joint_parent; M.joint_q_start = d->joint_q_start;
  M.joint_qd_start = d->joint_qd_start; M.joint_X_p = d->joint_X_p; M.joint_X_c = d->joint_X_c;
  M.joint_axis = d->joint_axis; M.joint_target_pos = d->joint_target_pos; M.joint_target_ke = d->joint_target_ke;
  M.joint_target_kd = d->joint_target_kd; M.joint_armature = d->joint_armature;
  M.mj_solref = d->mj_solref; M.mj_solimp = d->mj_solimp; M.mj_solmix = d->mj_solmix;
  M.mj_priority = d->mj_priority; M.mj_condim = d->mj_condim; M.mj_zero_margins = d->mj_zero_margins ? 1 : 0;
  m->has_mj = d->mj_solref && d->mj_solimp && d->mj_solmix && d->mj_priority && d->mj_condim;
  if (m->has_mj && M.S > 0) {
    const int S = M.S;
    std::vector<int32_t> cd(S);
    if (hipMemcpy(cd.data(), d->mj_condim, sizeof(int32_t) * S, hipMemcpyDeviceToHost) != hipSuccess) {
      delete m;
      return fail("na_model_create: could not read mj_condim (device pointer?)");
    }
    for (int i = 0; i < S; ++i)
      if (cd[i] != 1 && cd[i] != 3) {
        delete m;
        return fail("na_model_create: mj_condim must be 1 or 3 (torsional and rolling friction are not modelled)");
      }
  }
  if (m->articulated) {
    err = validate_articulation(M, wbs, wbe);
    if (!err.empty()) { delete m; return fail(err.c_str()); }
  }
  if (m->max_nb > NA_MAX_BODIES) { delete m; return fail("na_model_create: a world has more than 64 bodies"); }
  if (m->max_seg > NA_MAX_SEG) {
    delete m;
    return fail("na_model_create: a world segment (world shapes + global shapes) exceeds 128 shapes");
  }
  *out = m;
  return 0;
}
extern "C" int na_model_destroy(na_model* m) {
  delete m;
  return 0;
}

// broad phase code valid for this model (explicit mode needs the model's pair list)
static bool broad_phase_ok(const na_model* m, int bp) {
  return bp == 0 || bp == 1 || (bp == 2 && m->M.P > 0);
}

static NaControl to_control(const na_control_desc* c) {
  NaControl u{};
  if (c) {
    u.jf = c->joint_f;
    u.tq = c->joint_target_q;
    u.tqd = c->joint_target_qd;
    u.act = c->joint_act;
  }
  return u;
}

static NaConfig to_cfg(const na_config* c) {
  NaConfig k;
  k.mode = c->mode; k.controller = c->controller; k.tol = c->tol; k.dt_init = c->dt_inner_init;
  k.dt_min = c->dt_inner_min; k.dt_max = c->dt_inner_max; k.divergence_threshold = c->divergence_threshold;
  k.max_substeps = c->max_substeps; k.max_iterations = c->max_iterations;
  k.max_contacts = c->max_rigid_contact < NA_MAX_CONTACTS ? c->max_rigid_contact : NA_MAX_CONTACTS;
  k.broad_phase = c->broad_phase; k.opt_rel_tol = c->optimality_rel_tol; k.opt_abs_tol = c->optimality_abs_tol;
  k.tau_d = c->contact_tau_d; k.sigma = c->sigma; k.beta = c->beta; k.ls_rel_tol = c->ls_rel_tol;
  k.cg_rel_tol = c->cg_rel_tol; k.max_cg = c->max_cg; k.cg_forcing = c->cg_forcing;
  k.line_search = c->line_search;
  k.contact_law = c->contact_law;
  k.impratio = c->impratio;
  k.warm_start = c->warm_start;
  return k;
}

extern "C" int na_solver_create(const na_model* m, const na_config* c, const na_telemetry* buffers,
                                na_solver** out) {
  if (!m || !c || !out) return fail("na_solver_create: null argument");
  if (c->mode != 0 && c->mode != 1) return fail("na_solver_create: mode must be 0 (fixed) or 1 (adaptive)");
  if (c->controller != 0 && c->controller != 1) return fail("na_solver_create: controller must be 0 or 1");
  if (!(c->tol > 0)) return fail("na_solver_create: tol must be > 0");
  if (!(c->dt_inner_init > 0)) return fail("na_solver_create: dt_inner_init must be > 0");
  if (!(c->dt_inner_min > 0)) return fail("na_solver_create: dt_inner_min must be > 0");
  if (!(c->dt_inner_max > 0)) return fail("na_solver_create: dt_inner_max must be > 0");
  if (c->max_substeps < 1) return fail("na_solver_create: max_substeps must be >= 1");
  if (c->max_rigid_contact < 1 || c->max_rigid_contact > NA_MAX_CONTACTS)
    return fail("na_solver_create: max_rigid_contact must be in [1, 128] per world");
  if (!broad_phase_ok(m, c->broad_phase))
    return fail("na_solver_create: broad_phase must be 0 (sap), 1 (nxn) or 2 (explicit, needs the model's pairs)");
  if (c->line_search != 0 && c->line_search != 1)
    return fail("na_solver_create: line_search must be 0 (exact_root) or 1 (armijo_decay)");
  if (c->contact_law != 0 && c->contact_law != 1)
    return fail("na_solver_create: contact_law must be 0 (SAP) or 1 (MuJoCo soft constraints)");
  if (c->contact_law == 1) {
    if (!m->has_mj) return fail("na_solver_create: contact_law 1 needs the model's mj_* shape arrays");
    if (m->articulated)
      return fail("na_solver_create: contact_law 1 (MuJoCo soft constraints) is built for free-body models only");
    if (!(c->impratio > 0)) return fail("na_solver_create: impratio must be > 0");
  }
  if (c->warm_start < 0 || c->warm_start > 2)
    return fail("na_solver_create: warm_start must be 0 (reference seeds), 1 (extrapolated seeds) or 2 "
                "(concurrent seeds)");
  if (c->warm_start >= 1 && buffers && !buffers->body_trend)
    return fail("na_solver_create: warm_start 1 or 2 with supplied buffers needs body_trend");
  na_solver* s = new na_solver();
  s->model = m;
  s->cfg = to_cfg(c);
  s->pool = nullptr;
  int W = m->M.W, B = m->M.B;
  char* pool = nullptr;
  size_t off = 0;
  auto take = [&](size_t n) {
    char* p = pool + off;
    off += (n + 255) & ~(size_t)255;
    return p;
  };
  size_t need = 0;
  size_t sizes_own[] = {8u * W, 8u * W, 8u * W, 8u * W, 8u * W, 8u * W, 8u * W, (size_t)W, (size_t)W, (size_t)W,
                        4u * W, 4u * W, 4u * W, 4u * W, 8u, sizeof(long long) * NA_WORK * W,
                        sizeof(double) * 7 * (size_t)(B > 0 ? B : 1), 4u, 4u,
                        sizeof(float) * 6 * (size_t)(B > 0 ? B : 1)};
  size_t sizes_ext[] = {4u, 4u};
  if (buffers) for (size_t z : sizes_ext) need += (z + 255) & ~(size_t)255;
  else for (size_t z : sizes_own) need += (z + 255) & ~(size_t)255;
  if (hipMalloc(&pool, need) != hipSuccess) { delete s; return fail("na_solver_create: hipMalloc failed"); }
  if (hipMemset(pool, 0, need) != hipSuccess) { (void)hipFree(pool); delete s; return fail("na_solver_create: memset"); }
  s->pool = pool;
  Ctrl& K = s->K;
  if (buffers) {
    const na_telemetry* b = buffers;
    if (!b->dt || !b->dt_half || !b->ideal_dt || !b->sim_time || !b->next_time || !b->last_raw_error ||
        !b->last_error || !b->accepted || !b->commit || !b->diverged || !b->substeps || !b->iterations ||
        !b->solve_fail || !b->overflow || !b->cum_accepted || !b->work || !b->body_q64) {
      (void)hipFree(pool);
      delete s;
      return fail("na_solver_create: every buffer pointer must be set when buffers are supplied");
    }
    K.dt = b->dt; K.dt_half = b->dt_half; K.ideal_dt = b->ideal_dt; K.sim_time = b->sim_time;
    K.next_time = b->next_time; K.last_error = b->last_raw_error; K.acc_error = b->last_error;
    K.accepted = b->accepted; K.commit = b->commit; K.diverged = b->diverged; K.substeps = b->substeps;
    K.iterations = b->iterations; K.solve_fail = b->solve_fail; K.overflow = b->overflow;
    K.cum = b->cum_accepted; K.work = b->work; s->q64 = b->body_q64;
    K.trend = c->warm_start >= 1 ? b->body_trend : nullptr;
  } else {
    K.dt = (double*)take(8u * W); K.dt_half = (double*)take(8u * W); K.ideal_dt = (double*)take(8u * W);
    K.sim_time = (double*)take(8u * W); K.next_time = (double*)take(8u * W); K.last_error = (double*)take(8u * W);
    K.acc_error = (double*)take(8u * W);
    K.accepted = (uint8_t*)take(W); K.commit = (uint8_t*)take(W); K.diverged = (uint8_t*)take(W);
    K.substeps = (int32_t*)take(4u * W); K.iterations = (int32_t*)take(4u * W);
    K.solve_fail = (int32_t*)take(4u * W); K.overflow = (int32_t*)take(4u * W);
    K.cum = (unsigned long long*)take(8);
    K.work = (long long*)take(sizeof(long long) * NA_WORK * W);
    s->q64 = (double*)take(sizeof(double) * 7 * (size_t)(B > 0 ? B : 1));
  }
  s->queue = (int32_t*)take(8);  // [0] the list counter; [1] the second list of k_serve_mixed_tile
  s->flag = (int32_t*)take(4);
  if (!buffers) {
    float* tr = (float*)take(sizeof(float) * 6 * (size_t)(B > 0 ? B : 1));
    K.trend = c->warm_start >= 1 ? tr : nullptr;
  }
  if (hipHostMalloc(&s->flag_host, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(pool);
    delete s;
    return fail("na_solver_create: hipHostMalloc failed");
  }
  s->order = nullptr;
  s->grid_waves = 0;
  s->tiles = 0;
  s->plain_batch = 0;
  // controller defaults (reset of all worlds)
  k_reset<<<blocks(W, 256), 256>>>(m->M, nullptr, s->cfg.dt_init, K);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  *out = s;
  return 0;
}
extern "C" int na_solver_destroy(na_solver* s) {
  if (!s) return 0;
  if (s->pool) (void)hipFree(s->pool);
  if (s->flag_host) (void)hipHostFree(s->flag_host);
  delete s;
  return 0;
}
extern "C" int na_solver_telemetry(const na_solver* s, na_telemetry* t) {
  if (!s || !t) return fail("na_solver_telemetry: null argument");
  const Ctrl& K = s->K;
  t->dt = K.dt; t->dt_half = K.dt_half; t->ideal_dt = K.ideal_dt; t->sim_time = K.sim_time;
  t->next_time = K.next_time; t->last_raw_error = K.last_error; t->last_error = K.acc_error;
  t->accepted = K.accepted; t->commit = K.commit; t->diverged = K.diverged; t->substeps = K.substeps;
  t->iterations = K.iterations; t->solve_fail = K.solve_fail; t->overflow = K.overflow;
  t->cum_accepted = K.cum; t->work = K.work; t->body_q64 = s->q64; t->body_trend = K.trend;
  return 0;
}
extern "C" int na_solver_set_schedule(na_solver* s, int grid_waves, const int32_t* order) {
  if (!s) return fail("na_solver_set_schedule: null solver");
  s->grid_waves = grid_waves;
  s->order = order;
  return 0;
}

extern "C" int na_solver_set_tiles(na_solver* s, int32_t tiles) {
  if (!s) return fail("na_solver_set_tiles: null solver");
  if (tiles < 0) return fail("na_solver_set_tiles: tiles must be >= 0");
  if (tiles > 0 && (s->cfg.mode != 1 || s->cfg.warm_start != 2 || s->cfg.contact_law != 0))
    return fail("na_solver_set_tiles: two-wave tiles need adaptive mode, warm_start 2 (concurrent seeds) and the "
                "SAP contact law");
  s->tiles = tiles;
  return 0;
}

extern "C" int na_solver_set_plain_batch(na_solver* s, int32_t batch_all) {
  if (!s) return fail("na_solver_set_plain_batch: null solver");
  s->plain_batch = batch_all ? 1 : 0;
  return 0;
}

// Grid of a tile launch over n listed worlds, workgroups of `waves` waves: as many workgroups as
// fit (or the scheduled waves), no more than the list needs to give every world its waves at
// once: a tile takes two waves, any other world one, so (n + tiles) / waves workgroups, and at
// least one per tile.
static int tile_grid(const na_solver* s, const void* kernel, int n, int waves) {
  int dev = 0, cus = 256, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * waves, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  int grid = cus * per_cu;
  if (s->grid_waves > 0) grid = (s->grid_waves + waves - 1) / waves;
  const int tiles = s->tiles < n ? s->tiles : n;
  int need = (n + tiles + waves - 1) / waves;
  need = need > tiles ? need : tiles;
  grid = grid < need ? grid : need;
  return grid > 0 ? grid : 1;
}

static int default_grid(const na_solver* s, const void* kernel = (const void*)k_frame) {
  int W = s->model->M.W;
  if (s->grid_waves > 0) return s->grid_waves < W ? s->grid_waves : W;
  int dev = 0, cus = 256, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  int g = cus * per_cu;
  return g < W ? g : W;
}

extern "C" int na_step(na_solver* s, float* body_q, float* body_qd, const float* body_f,
                       const na_control_desc* control, double dt_outer, void* stream) {
  if (!s || !body_q || !body_qd) return fail("na_step: null argument");
  if (!(dt_outer > 0)) return fail("na_step: dt (dt_outer) must be > 0");
  const NaModel& M = s->model->M;
  double eff = s->cfg.dt_max < dt_outer ? s->cfg.dt_max : dt_outer;
  k_zero_i32<<<1, 64, 0, S_(stream)>>>(s->queue, 1);
  if (s->tiles > 0) {
    int grid = tile_grid(s, (const void*)k_frame_tile, M.W, 2);
    k_frame_tile<<<grid, 128, 0, S_(stream)>>>(M, s->cfg, s->K, s->q64, body_q, body_qd, body_f,
                                               to_control(control), dt_outer, eff, s->queue, s->order, s->tiles);
    HIPCHK(hipGetLastError());
    return 0;
  }
  int grid = default_grid(s);
  k_frame<<<grid, 64, 0, S_(stream)>>>(M, s->cfg, s->K, s->q64, body_q, body_qd, body_f, to_control(control),
                                       dt_outer, eff, s->queue, s->order);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_serve(na_solver* s, const int32_t* active, int32_t n_active, const int32_t* horizon_frames,
                        const int32_t* horizon_accepted, int32_t* frames_done, int32_t* accepted, int32_t* status,
                        const double* world_tol, float* body_q, float* body_qd, const float* body_f,
                        double dt_outer, double budget_ms, int32_t max_iterations, int32_t* unfinished,
                        void* stream) {
  if (!s || !body_q || !body_qd) return fail("na_serve: null argument");
  if (n_active < 0 || n_active > s->model->M.W) return fail("na_serve: n_active must be in [0, world_count]");
  if (n_active > 0 && (!active || !horizon_frames || !horizon_accepted || !frames_done || !accepted || !status))
    return fail("na_serve: null slot array");
  if (!(dt_outer > 0)) return fail("na_serve: dt_outer must be > 0");
  if (n_active == 0) return 0;
  const NaModel& M = s->model->M;
  hipStream_t st = S_(stream);
  int dev = 0, rate_khz = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
  long long budget = budget_ms > 0 ? (long long)(budget_ms * (double)rate_khz) : (1LL << 62);
  double eff = s->cfg.dt_max < dt_outer ? s->cfg.dt_max : dt_outer;
  ServeSlots SV{horizon_frames, horizon_accepted, frames_done, accepted, status, world_tol};
  k_zero_i32<<<1, 64, 0, st>>>(s->queue, 1);
  if (unfinished) k_zero_i32<<<1, 64, 0, st>>>(unfinished, 1);
  if (s->tiles > 0) {
    int grid = tile_grid(s, (const void*)k_serve_tile, n_active, 2);
    k_serve_tile<<<grid, 128, 0, st>>>(M, s->cfg, s->K, s->q64, body_q, body_qd, body_f, dt_outer, eff, s->queue,
                                       active, n_active, s->tiles, SV, budget, max_iterations, unfinished);
    HIPCHK(hipGetLastError());
    return 0;
  }
  int grid = default_grid(s);
  grid = grid < n_active ? grid : n_active;
  k_serve<<<grid, 64, 0, st>>>(M, s->cfg, s->K, s->q64, body_q, body_qd, body_f, dt_outer, eff, s->queue, active,
                               n_active, SV, budget, max_iterations, unfinished);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_step_plain(na_solver* s, float* body_q, float* body_qd, const float* body_f,
                             const na_control_desc* control, double dt_outer, int32_t* iterations_out, void* stream) {
  if (!s || !body_q || !body_qd) return fail("na_step_plain: null argument");
  if (!(dt_outer > 0)) return fail("na_step_plain: dt (dt_outer) must be > 0");
  const NaModel& M = s->model->M;
  hipStream_t st = S_(stream);
  double eff = s->cfg.dt_max < dt_outer ? s->cfg.dt_max : dt_outer;
  k_sync_shadow<<<blocks(M.B, 256), 256, 0, st>>>(M.B, s->q64, body_q);
  k_plain_open<<<blocks(M.W, 256), 256, 0, st>>>(M, s->cfg, s->K, dt_outer, eff);
  int status = 0, it = 0;
  for (; it < s->cfg.max_substeps; ++it) {
    k_zero_i32<<<1, 64, 0, st>>>(s->flag, 1);
    k_plain_iter<<<M.W, 64, 0, st>>>(M, s->cfg, s->K, s->q64, body_qd, body_f, to_control(control), eff, s->flag,
                                     s->plain_batch);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s->flag_host, s->flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int f = *s->flag_host;
    if (f >= 2) { status = 2; ++it; break; }
    if (f == 0) { ++it; break; }
  }
  if (status == 0 && it >= s->cfg.max_substeps && *s->flag_host != 0) status = 1;
  k_round_shadow<<<blocks(7LL * M.B, 256), 256, 0, st>>>(M.B, s->q64, body_q);
  HIPCHK(hipGetLastError());
  if (iterations_out) *iterations_out = it;
  return status;
}

extern "C" int na_coords_to_mujoco(const na_model* m, const float* joint_q, const float* joint_qd, float* qpos,
                                   float* qvel, void* stream) {
  if (!m || !joint_q || !joint_qd || !qpos || !qvel) return fail("na_coords_to_mujoco: null argument");
  const NaModel& M = m->M;
  int B = M.B;
  if (B == 0) return 0;
  if (m->articulated)
    k_coords_to_mujoco_art<<<blocks(B, 256), 256, 0, S_(stream)>>>(B, M.joint_type, M.joint_q_start, M.joint_qd_start,
                                                                   M.joint_X_p, M.body_com, joint_q, joint_qd, qpos,
                                                                   qvel);
  else
    k_coords_to_mujoco<<<blocks(B, 256), 256, 0, S_(stream)>>>(B, M.body_com, joint_q, joint_qd, qpos, qvel);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_coords_from_mujoco(const na_model* m, const float* qpos, const float* qvel, float* joint_q,
                                     float* joint_qd, void* stream) {
  if (!m || !joint_q || !joint_qd || !qpos || !qvel) return fail("na_coords_from_mujoco: null argument");
  const NaModel& M = m->M;
  int B = M.B;
  if (B == 0) return 0;
  if (m->articulated)
    k_coords_from_mujoco_art<<<blocks(B, 256), 256, 0, S_(stream)>>>(B, M.joint_type, M.joint_q_start,
                                                                     M.joint_qd_start, M.joint_X_p, M.body_com, qpos,
                                                                     qvel, joint_q, joint_qd);
  else
    k_coords_from_mujoco<<<blocks(B, 256), 256, 0, S_(stream)>>>(B, M.body_com, qpos, qvel, joint_q, joint_qd);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_reset(na_solver* s, const uint8_t* mask, void* stream) {
  if (!s) return fail("na_reset: null solver");
  int W = s->model->M.W;
  k_reset<<<blocks(W, 256), 256, 0, S_(stream)>>>(s->model->M, mask, s->cfg.dt_init, s->K);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int64_t na_collide_scratch_words(const na_model* m, int32_t per_world) {
  if (!m) return -1;
  return (int64_t)m->M.W * per_world * NA_STAGE_WORDS + 2LL * m->M.W;
}

extern "C" int na_collide(const na_model* m, const float* body_q, const na_contacts_desc* out, int32_t per_world,
                          int32_t broad_phase, void* stream) {
  if (!m || !body_q || !out) return fail("na_collide: null argument");
  if (per_world < 1 || per_world > NA_MAX_CONTACTS) return fail("na_collide: per-world capacity must be in [1, 128]");
  if (!broad_phase_ok(m, broad_phase)) return fail("na_collide: invalid broad_phase for this model");
  const NaModel& M = m->M;
  hipStream_t st = S_(stream);
  NaConfig cfg{};
  cfg.max_contacts = per_world;
  cfg.broad_phase = broad_phase;
  int32_t* staging = out->scratch;
  int32_t* offsets = out->scratch + (long long)M.W * per_world * NA_STAGE_WORDS;
  k_collide_api<<<M.W, 64, 0, st>>>(M, cfg, body_q, per_world, staging, out->world_count_out);
  k_scan_counts<<<1, 1024, 0, st>>>(M.W, per_world, out->world_count_out, offsets, out->count);
  k_compact_contacts<<<M.W, 64, 0, st>>>(M.W, per_world, staging, out->world_count_out, offsets, *out);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- allocation-free contact generation
// na_collide_noalloc keeps na_collide's staging area and per-world offsets in this file-scope
// array instead of the caller's scratch.  Each GPU holds its own copy for the life of the process;
// host code reaches it only through the per-device address from hipGetSymbolAddress (the array's
// name is a host-side stand-in, not a device address).
static __device__ __align__(256) int32_t scratch[NA_COLLIDE_NOALLOC_MAX_WORDS];
#define NA_MAX_DEVICES 64
static int32_t* g_scratch_dev[NA_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per device.
static int32_t* scratch_on_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_dev[dev] = (int32_t*)p;
  }
  return g_scratch_dev[dev];
}

extern "C" int na_collide_noalloc(const na_model* m, const float* body_q, const na_contacts_desc* out,
                                  int32_t per_world, int32_t broad_phase, void* stream) {
  if (!m || !body_q || !out || !out->count || !out->world_count_out)
    return fail("na_collide_noalloc: null argument");
  if (per_world < 1 || per_world > NA_MAX_CONTACTS)
    return fail("na_collide_noalloc: per-world capacity must be in [1, 128]");
  if (!broad_phase_ok(m, broad_phase)) return fail("na_collide_noalloc: invalid broad_phase for this model");
  const NaModel& M = m->M;
  const long long words = (long long)M.W * per_world * NA_STAGE_WORDS + 2LL * M.W;
  if (words > (long long)NA_COLLIDE_NOALLOC_MAX_WORDS)
    return fail("na_collide_noalloc: world_count * per_world * 20 + 2 * world_count exceeds the static "
                "scratch (NA_COLLIDE_NOALLOC_MAX_WORDS words); use na_collide with a caller scratch");
  int32_t* base = scratch_on_current_device();
  if (!base) return fail("na_collide_noalloc: hipGetSymbolAddress of the static scratch failed");
  hipStream_t st = S_(stream);
  NaConfig cfg{};
  cfg.max_contacts = per_world;
  cfg.broad_phase = broad_phase;
  // every word the kernels read is written first in this call (staging slots below each world's
  // count, all W offsets), so nothing is cleared
  int32_t* staging = base;
  int32_t* offsets = base + (long long)M.W * per_world * NA_STAGE_WORDS;
  k_collide_api<<<M.W, 64, 0, st>>>(M, cfg, body_q, per_world, staging, out->world_count_out);
  k_scan_counts<<<1, 1024, 0, st>>>(M.W, per_world, out->world_count_out, offsets, out->count);
  k_compact_contacts<<<M.W, 64, 0, st>>>(M.W, per_world, staging, out->world_count_out, offsets, *out);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_match_contacts(const na_model* m, const float* body_q, const na_contacts_desc* cur,
                                 int32_t per_world, const int32_t* cur_offsets, const int32_t* prev_shape0,
                                 const int32_t* prev_shape1, const float* prev_mid, const float* prev_normal,
                                 const int32_t* prev_offsets, const int32_t* prev_counts, int32_t prev_capacity,
                                 float pos_threshold, float normal_dot_threshold, int32_t* match_index,
                                 float* cur_mid, int32_t* prev_matched, const float* prev_point0,
                                 const float* prev_point1, const float* prev_offset0, const float* prev_offset1,
                                 void* stream) {
  if (!m || !body_q || !cur || !cur_offsets || !prev_shape0 || !prev_shape1 || !prev_mid || !prev_normal ||
      !prev_offsets || !prev_counts || !match_index || !cur_mid || !prev_matched)
    return fail("na_match_contacts: null argument");
  if (per_world < 1 || per_world > NA_MAX_CONTACTS) return fail("na_match_contacts: per-world capacity in [1, 128]");
  if (!(pos_threshold >= 0.f) || !(normal_dot_threshold <= 1.f)) return fail("na_match_contacts: bad thresholds");
  const int sticky = (prev_point0 != nullptr) + (prev_point1 != nullptr) + (prev_offset0 != nullptr) +
                     (prev_offset1 != nullptr);
  if (sticky != 0 && sticky != 4) return fail("na_match_contacts: sticky mode needs all four previous geometry arrays");
  const NaModel& M = m->M;
  k_match_contacts<<<M.W, 64, 0, S_(stream)>>>(M, body_q, *cur, per_world, cur_offsets, prev_shape0, prev_shape1,
                                               prev_mid, prev_normal, prev_offsets, prev_counts, prev_capacity,
                                               pos_threshold * pos_threshold, normal_dot_threshold, match_index,
                                               cur_mid, prev_matched, prev_point0, prev_point1, prev_offset0,
                                               prev_offset1);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_collide_debug(const na_model* m, const float* body_q, float* lo, float* hi, int32_t* pairs,
                                int32_t* pair_count, int32_t max_pairs, int32_t broad_phase, void* stream) {
  if (!m || !body_q || !lo || !hi || !pairs || !pair_count) return fail("na_collide_debug: null argument");
  if (!broad_phase_ok(m, broad_phase)) return fail("na_collide_debug: invalid broad_phase for this model");
  const NaModel& M = m->M;
  k_collide_debug<<<M.W, 64, 0, S_(stream)>>>(M, broad_phase, body_q, lo, hi, pairs, pair_count, max_pairs);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_substep(const na_solver* s, const double* q64, const float* qd, const float* f, const double* h,
                          const double* guess, double* q_out, float* qd_out, double* v_solution, int32_t* stats,
                          void* stream) {
  if (!s || !q64 || !qd || !h || !q_out || !qd_out || !stats) return fail("na_substep: null argument");
  const NaModel& M = s->model->M;
  k_substep<<<M.W, 64, 0, S_(stream)>>>(M, s->cfg, q64, qd, f, h, guess, q_out, qd_out, v_solution, stats);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int na_attempt(const na_solver* s, const double* q64, const float* qd, const float* f, const double* dt,
                          double* q_full, float* qd_full, double* q_dbl, float* qd_dbl, double* err,
                          int32_t* solve_ok, int32_t* n_contacts, float* contacts_out, int32_t* contact_counts_out,
                          const float* trend, void* stream) {
  if (!s || !q64 || !qd || !dt || !q_full || !qd_full || !q_dbl || !qd_dbl || !err || !solve_ok || !n_contacts)
    return fail("na_attempt: null argument");
  if ((contacts_out == nullptr) != (contact_counts_out == nullptr))
    return fail("na_attempt: contacts_out and contact_counts_out go together");
  const NaModel& M = s->model->M;
  k_attempt<<<M.W, 64, 0, S_(stream)>>>(M, s->cfg, q64, qd, f, dt, q_full, qd_full, q_dbl, qd_dbl, err, solve_ok,
                                        n_contacts, contacts_out, contact_counts_out,
                                        s->cfg.warm_start >= 1 ? trend : nullptr);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- articulated entry points
extern "C" int na_step_articulated(na_solver* s, float* joint_q, float* joint_qd, float* body_q, float* body_qd,
                                   const float* body_f, const na_control_desc* control, double dt_outer,
                                   void* stream) {
  if (!s || !joint_q || !joint_qd) return fail("na_step_articulated: null argument");
  if (!s->model->articulated) return fail("na_step_articulated: the model has no joints");
  if (!(dt_outer > 0)) return fail("na_step_articulated: dt (dt_outer) must be > 0");
  const NaModel& M = s->model->M;
  const double eff = s->cfg.dt_max < dt_outer ? s->cfg.dt_max : dt_outer;
  k_zero_i32<<<1, 64, 0, S_(stream)>>>(s->queue, 1);
  const int grid = default_grid(s, (const void*)k_art_frame);
  k_art_frame<<<grid, 64, 0, S_(stream)>>>(M, s->cfg, s->K, s->q64, joint_q, joint_qd, body_q, body_qd, body_f,
                                           to_control(control), dt_outer, eff, s->queue, s->order);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_serve_articulated(na_solver* s, const int32_t* active, int32_t n_active,
                                    const int32_t* horizon_frames, const int32_t* horizon_accepted,
                                    int32_t* frames_done, int32_t* accepted, int32_t* status, const double* world_tol,
                                    float* joint_q, float* joint_qd, double* joint_q64, double* joint_qd64,
                                    float* body_q, float* body_qd, double dt_outer, double budget_ms,
                                    int32_t max_iterations, int32_t* unfinished, void* stream) {
  if (!s || !joint_q || !joint_qd || !joint_q64 || !joint_qd64) return fail("na_serve_articulated: null argument");
  if (!s->model->articulated) return fail("na_serve_articulated: the model has no joints");
  if (n_active < 0 || n_active > s->model->M.W) return fail("na_serve_articulated: n_active must be in [0, world_count]");
  if (n_active > 0 && (!active || !horizon_frames || !horizon_accepted || !frames_done || !accepted || !status))
    return fail("na_serve_articulated: null slot array");
  if (!(dt_outer > 0)) return fail("na_serve_articulated: dt_outer must be > 0");
  if (n_active == 0) return 0;
  const NaModel& M = s->model->M;
  hipStream_t st = S_(stream);
  int dev = 0, rate_khz = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
  const long long budget = budget_ms > 0 ? (long long)(budget_ms * (double)rate_khz) : (1LL << 62);
  const double eff = s->cfg.dt_max < dt_outer ? s->cfg.dt_max : dt_outer;
  ServeSlots SV{horizon_frames, horizon_accepted, frames_done, accepted, status, world_tol};
  k_zero_i32<<<1, 64, 0, st>>>(s->queue, 1);
  int grid = default_grid(s, (const void*)k_art_serve);
  grid = grid < n_active ? grid : n_active;
  if (unfinished) k_zero_i32<<<1, 64, 0, st>>>(unfinished, 1);
  ArtServeState X{joint_q64, joint_qd64, joint_q, joint_qd, body_q, body_qd, nullptr};
  k_art_serve<<<grid, 64, 0, st>>>(M, s->cfg, s->K, X, dt_outer, eff, s->queue, active, n_active, SV, budget,
                                   max_iterations, unfinished);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_serve_mixed(na_solver* s, const int32_t* active, int32_t n_active, const int32_t* horizon_frames,
                              const int32_t* horizon_accepted, int32_t* frames_done, int32_t* accepted,
                              int32_t* status, const double* world_tol, float* body_q, float* body_qd,
                              const float* body_f, float* joint_q, float* joint_qd, double* joint_q64,
                              double* joint_qd64, double dt_outer, double budget_ms, int32_t max_iterations,
                              int32_t* unfinished, void* stream) {
  if (!s || !body_q || !body_qd || !joint_q || !joint_qd || !joint_q64 || !joint_qd64)
    return fail("na_serve_mixed: null argument");
  if (!s->model->articulated) return fail("na_serve_mixed: the model has no joint arrays");
  if (n_active < 0 || n_active > s->model->M.W) return fail("na_serve_mixed: n_active must be in [0, world_count]");
  if (n_active > 0 && (!active || !horizon_frames || !horizon_accepted || !frames_done || !accepted || !status))
    return fail("na_serve_mixed: null slot array");
  if (!(dt_outer > 0)) return fail("na_serve_mixed: dt_outer must be > 0");
  if (n_active == 0) return 0;
  const NaModel& M = s->model->M;
  hipStream_t st = S_(stream);
  int dev = 0, rate_khz = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
  const long long budget = budget_ms > 0 ? (long long)(budget_ms * (double)rate_khz) : (1LL << 62);
  const double eff = s->cfg.dt_max < dt_outer ? s->cfg.dt_max : dt_outer;
  ServeSlots SV{horizon_frames, horizon_accepted, frames_done, accepted, status, world_tol};
  ArtServeState X{joint_q64, joint_qd64, joint_q, joint_qd, body_q, body_qd, body_f};
  if (unfinished) k_zero_i32<<<1, 64, 0, st>>>(unfinished, 1);
  if (s->tiles > 0) {
    k_zero_i32<<<1, 64, 0, st>>>(s->queue, 2);
    const int grid = tile_grid(s, (const void*)k_serve_mixed_tile, n_active, 3);
    k_serve_mixed_tile<<<grid, 192, 0, st>>>(M, s->cfg, s->K, s->q64, body_q, body_qd, body_f, X, dt_outer, eff,
                                             s->queue, active, n_active, s->tiles, SV, budget, max_iterations,
                                             unfinished);
    HIPCHK(hipGetLastError());
    return 0;
  }
  k_zero_i32<<<1, 64, 0, st>>>(s->queue, 1);
  int grid = default_grid(s, (const void*)k_serve_mixed);
  grid = grid < n_active ? grid : n_active;
  k_serve_mixed<<<grid, 64, 0, st>>>(M, s->cfg, s->K, s->q64, body_q, body_qd, body_f, X, dt_outer, eff, s->queue,
                                     active, n_active, SV, budget, max_iterations, unfinished);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_art_substep(const na_solver* s, const double* joint_q, const double* joint_qd, const double* h,
                              const float* body_f, const na_control_desc* control, double* q_out, double* qd_out,
                              int32_t* stats, void* stream) {
  if (!s || !joint_q || !joint_qd || !h || !q_out || !qd_out || !stats) return fail("na_art_substep: null argument");
  if (!s->model->articulated) return fail("na_art_substep: the model has no joints");
  const NaModel& M = s->model->M;
  k_art_substep<<<M.W, 64, 0, S_(stream)>>>(M, s->cfg, joint_q, joint_qd, h, body_f, to_control(control), q_out,
                                             qd_out, stats);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_art_attempt(const na_solver* s, const double* joint_q, const double* joint_qd, const double* dt,
                              const float* body_f, const na_control_desc* control, double* q_full, double* qd_full,
                              double* q_dbl, double* qd_dbl, double* err, int32_t* solve_ok, int32_t* n_contacts,
                              float* contacts_out, int32_t* contact_counts_out, void* stream) {
  if (!s || !joint_q || !joint_qd || !dt || !q_full || !qd_full || !q_dbl || !qd_dbl || !err || !solve_ok ||
      !n_contacts)
    return fail("na_art_attempt: null argument");
  if ((contacts_out == nullptr) != (contact_counts_out == nullptr))
    return fail("na_art_attempt: contacts_out and contact_counts_out go together");
  if (!s->model->articulated) return fail("na_art_attempt: the model has no joints");
  const NaModel& M = s->model->M;
  k_art_attempt<<<M.W, 64, 0, S_(stream)>>>(M, s->cfg, joint_q, joint_qd, dt, body_f, to_control(control), q_full,
                                             qd_full, q_dbl, qd_dbl, err, solve_ok, n_contacts, contacts_out,
                                             contact_counts_out);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- controller entry points
extern "C" int na_ctrl_calc_adjusted_step_f32(int32_t n, const float* err, const float* dt, float* ideal,
                                              uint8_t* acc, uint8_t* com, uint8_t* div, float tol, float dt_min,
                                              float thr, void* stream) {
  if (n <= 0) return 0;
  k_calc_adjusted_step_f32<<<blocks(n, 256), 256, 0, S_(stream)>>>(n, err, dt, ideal, acc, com, div, tol, dt_min, thr);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_ctrl_adapt_dt_f32(int32_t n, const float* err, float* sim, const float* nxt, float* dt, float* dth,
                                    float* ideal, uint8_t* div, uint8_t* acc, float* acc_err, int32_t* sub,
                                    int32_t* cum, int32_t mode, float tol, float dt_min, float dt_max, float thr,
                                    void* stream) {
  if (n <= 0) return 0;
  k_adapt_dt_f32<<<blocks(n, 256), 256, 0, S_(stream)>>>(n, err, sim, nxt, dt, dth, ideal, div, acc, acc_err, sub, cum,
                                                         mode, tol, dt_min, dt_max, thr);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_ctrl_rebase_time_f32(int32_t n, float* sim, float* nxt, void* stream) {
  if (n <= 0) return 0;
  k_rebase_time_f32<<<blocks(n, 256), 256, 0, S_(stream)>>>(n, sim, nxt);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_ctrl_reset_worlds_f32(int32_t n, const uint8_t* mask, float dt_init, float* ideal, float* dt,
                                        float* dth, float* sim, float* nxt, uint8_t* div, uint8_t* acc, void* stream) {
  if (n <= 0) return 0;
  k_reset_worlds_f32<<<blocks(n, 256), 256, 0, S_(stream)>>>(n, mask, dt_init, ideal, dt, dth, sim, nxt, div, acc);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_ctrl_clamp_dt_to_boundary_f32(int32_t n, float* dt, float* dth, const float* sim, const float* nxt,
                                                void* stream) {
  if (n <= 0) return 0;
  k_clamp_dt_f32<<<blocks(n, 256), 256, 0, S_(stream)>>>(n, dt, dth, sim, nxt);
  HIPCHK(hipGetLastError());
  return 0;
}
extern "C" int na_ctrl_inf_norm_error_f32(int32_t n, int32_t coords, const float* qf, const float* qd,
                                          const float* scale, float* err, void* stream) {
  if (n <= 0) return 0;
  k_inf_norm_f32<<<blocks(n, 256), 256, 0, S_(stream)>>>(n, coords, qf, qd, scale, err);
  HIPCHK(hipGetLastError());
  return 0;
}
