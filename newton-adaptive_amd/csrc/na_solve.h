// Convex compliant-contact (SAP) solve of one world inside one wavefront.
//
// The problem is the one the reference solves through the external sap_warp solver
// (call sites solver_sap_adaptive.py:489-501,714-729): with A = diag(m I, I_world) per free
// body, free-motion velocity v* and contact Jacobian J,
//     l(v) = 1/2 |v - v*|_A^2 + sum_i 1/2 |P_F(y_i)|_R^2,  y_i = -R^-1 (J_i v - v_hat_i),
// minimised by Newton's method.  One lane owns one body (registers) and lanes stride over
// contacts (LDS).  The contact -> body sums (J^T gamma, J^T G J p, the blocks J^T G J) use no LDS
// atomics: contact lanes store one world-frame value per contact, and each body lane sums the
// entries of its own contacts, listed once per solve in contact order (build_entries).
// The Newton direction comes from block-Jacobi preconditioned conjugate gradients on
// H = A + J^T G J (matrix free), the step length from an exact line search seeded by the
// cubic Hermite minimiser (N&W eq. 3.59, test_adaptive_cubic_linesearch.py:36-60) or, as the
// "armijo_decay" variant, from backtracking with Armijo's condition.
// Optimality follows the SAP criterion |D g| <= eps_a + eps_r max(|D p|, |D J^T gamma|).
#pragma once
#include "na_collide.h"

struct NaConfig {
  int mode;        // 0 fixed, 1 adaptive
  int controller;  // 0 SAP (_adapt_dt), 1 MuJoCo (_calc_adjusted_step)
  double tol, dt_init, dt_min, dt_max, divergence_threshold;
  int max_substeps, max_iterations, max_contacts, broad_phase;  // broad_phase: 0 sap, 1 nxn
  float opt_rel_tol, opt_abs_tol, tau_d, sigma, beta, ls_rel_tol, cg_rel_tol;
  int max_cg;
  float cg_forcing;  // > 0: upper bound of the inexact-Newton PCG forcing term; 0: fixed cg_rel_tol
  int line_search;   // 0 exact (root of l'(alpha), cubic seed), 1 armijo_decay (backtracking)
  int contact_law;   // 0 SAP compliant cone, 1 MuJoCo soft constraints (pyramidal cone)
  float impratio;    // MuJoCo frictional-to-normal impedance ratio (contact law 1)
  int warm_start;    // 0 the reference's Newton seeds; 1 extrapolated seeds (see world_attempt)
};

// Body slot NA_DUMMY stands for the static side of a contact: its velocity entries stay zero
// for the whole solve, so contact loops gather from it without branching.
#define NA_DUMMY NA_MAX_BODIES
#define NA_BSLOT (NA_MAX_BODIES + 1)
struct BodyLDS {
  float q[7][NA_MAX_BODIES];     // f32 pose used for contact generation (SoA)
  float x[3][NA_MAX_BODIES];     // COM position at the step start
  float iwi[6][NA_MAX_BODIES];   // world inverse inertia (xx, yy, zz, xy, xz, yz)
  float im[NA_MAX_BODIES];       // inverse mass
  float v[6][NA_BSLOT];          // f32 exchange vector (PCG directions)
  double vd[6][NA_BSLOT];        // f64 exchange vector (Newton iterate, line-search direction)
  // Values that the body lanes sum, one phase at a time: per contact, the world-frame impulse of
  // J^T gamma (fd, FP64) and the world-frame Hessian block F^T G F (gw: xx yy zz xy xz yz); per
  // body entry, the wrench (force; torque about the body's COM) of the PCG product J^T G J p
  // (ew: written by the contact lane into its two entry slots, so the body lane's reads need no
  // index).
  union __attribute__((aligned(16))) {
    double fd[NA_MAX_CONTACTS][4];
    float gw[NA_MAX_CONTACTS][8];
    float ew[2 * NA_MAX_CONTACTS][8];
  } xb;
};

struct SolveLDS {
  // Body entries of the solve (build_entries): entry e of body b (e in [eoff[b], eoff[b] +
  // edeg[b])) is contact ent[e], with the lever arm from the body's COM and the side's sign
  // (-1 body A, +1 body B) in er[e].
  float er[2 * NA_MAX_CONTACTS][4] __attribute__((aligned(16)));
  uint8_t ent[2 * NA_MAX_CONTACTS];
  uint8_t eslot[NA_MAX_CONTACTS][2];  // the entries of contact c's sides A and B (dynamic sides only)
  uint16_t eoff[NA_MAX_BODIES], edeg[NA_MAX_BODIES];
  float t1[3][NA_MAX_CONTACTS];
  float ra[3][NA_MAX_CONTACTS];  // contact point - COM of body A
  float rb[3][NA_MAX_CONTACTS];
  float Rn[NA_MAX_CONTACTS], Rt[NA_MAX_CONTACTS], vhat[NA_MAX_CONTACTS];
  double iRn[NA_MAX_CONTACTS], iRt[NA_MAX_CONTACTS];  // FP64 reciprocals of Rn, Rt
};

// The segment scratch is only live during contact generation and the per-contact solver
// data only during a solve, so they share storage.
struct __attribute__((aligned(16))) WorldLDS {
  union {
    SegLDS seg;
    SolveLDS sol;
  };
  ContactLDS con;
  BodyLDS body;
  MjContactParams<NA_MAX_CONTACTS> mj;  // contact law 1 only
};

// Per-lane body registers for one solve.
struct BodyReg {
  float m, im;
  float I[6];    // world inertia sym (xx, yy, zz, xy, xz, yz)
  double vs[6];  // v*
  float D[6];    // diag(A)^-1/2
};

template <typename T>
__device__ __forceinline__ void sym_mul(const float* S, const T* x, T* y) {
  y[0] = S[0] * x[0] + S[3] * x[1] + S[4] * x[2];
  y[1] = S[3] * x[0] + S[1] * x[1] + S[5] * x[2];
  y[2] = S[4] * x[0] + S[5] * x[1] + S[2] * x[2];
}
template <typename T>
__device__ __forceinline__ void applyA(const BodyReg& b, const T* x, T* y) {
  y[0] = b.m * x[0];
  y[1] = b.m * x[1];
  y[2] = b.m * x[2];
  sym_mul<T>(b.I, x + 3, y + 3);
}

struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 dd(f3 a) { return d3{a.x, a.y, a.z}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// One contact as the solve sees it: frame rows t1, t2 = n x t1, n; lever arms; body slots.
// J_b row k = sign * (F_k, r x F_k), sign + for body B, - for body A.  An absent side (static
// shape, or a lane past the contact count) has slot NA_DUMMY and flag false.
struct CRow {
  f3 t1, t2, n, ra, rb;
  int sa, sb;
  int ea, eb;  // entry slots of the dynamic sides (build_entries)
  bool ha, hb;
};
__device__ __forceinline__ CRow load_row(const ContactLDS& C, const SolveLDS& X, int c, bool valid) {
  CRow r;
  r.n = mk(C.n[0][c], C.n[1][c], C.n[2][c]);
  r.t1 = mk(X.t1[0][c], X.t1[1][c], X.t1[2][c]);
  r.t2 = cross3(r.n, r.t1);
  r.ra = mk(X.ra[0][c], X.ra[1][c], X.ra[2][c]);
  r.rb = mk(X.rb[0][c], X.rb[1][c], X.rb[2][c]);
  int ba = C.ba[c], bb = C.bb[c];
  r.ha = valid && ba >= 0;
  r.hb = valid && bb >= 0;
  r.sa = r.ha ? ba : NA_DUMMY;
  r.sb = r.hb ? bb : NA_DUMMY;
  r.ea = X.eslot[c][0];
  r.eb = X.eslot[c][1];
  return r;
}
// relative contact velocity J v in the contact frame (f32, from B.v)
__device__ __forceinline__ f3 row_vel(const BodyLDS& B, const CRow& r) {
  f3 va = mk(B.v[0][r.sa], B.v[1][r.sa], B.v[2][r.sa]), wa = mk(B.v[3][r.sa], B.v[4][r.sa], B.v[5][r.sa]);
  f3 vb = mk(B.v[0][r.sb], B.v[1][r.sb], B.v[2][r.sb]), wb = mk(B.v[3][r.sb], B.v[4][r.sb], B.v[5][r.sb]);
  f3 rel = (vb + cross3(wb, r.rb)) - (va + cross3(wa, r.ra));
  return mk(dot3(r.t1, rel), dot3(r.t2, rel), dot3(r.n, rel));
}
// relative contact velocity J p in the contact frame (f32), with p held by the body lanes (lane b:
// body b's six components): each side's components come from its body's lane through the LDS
// crossbar (ds_bpermute), with no LDS store, fence or load; a static side reads zero.  Every lane
// of the wave must be active: a permute reads nothing from an inactive lane, so the permutes run
// unconditionally and the static sides are zeroed after them.
__device__ __forceinline__ f3 row_vel_lanes(const float* p, const CRow& r) {
  float a[6], b[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int pk = __float_as_int(p[k]);
    const float ta = __int_as_float(__builtin_amdgcn_ds_bpermute(r.sa << 2, pk));
    const float tb = __int_as_float(__builtin_amdgcn_ds_bpermute(r.sb << 2, pk));
    a[k] = r.ha ? ta : 0.f;
    b[k] = r.hb ? tb : 0.f;
  }
  f3 va = mk(a[0], a[1], a[2]), wa = mk(a[3], a[4], a[5]);
  f3 vb = mk(b[0], b[1], b[2]), wb = mk(b[3], b[4], b[5]);
  f3 rel = (vb + cross3(wb, r.rb)) - (va + cross3(wa, r.ra));
  return mk(dot3(r.t1, rel), dot3(r.t2, rel), dot3(r.n, rel));
}
// relative contact velocity J v in the contact frame (f64, from B.vd)
__device__ __forceinline__ d3 row_vel_d(const BodyLDS& B, const CRow& r) {
  d3 va = d3{B.vd[0][r.sa], B.vd[1][r.sa], B.vd[2][r.sa]}, wa = d3{B.vd[3][r.sa], B.vd[4][r.sa], B.vd[5][r.sa]};
  d3 vb = d3{B.vd[0][r.sb], B.vd[1][r.sb], B.vd[2][r.sb]}, wb = d3{B.vd[3][r.sb], B.vd[4][r.sb], B.vd[5][r.sb]};
  d3 xa = dcross(wa, dd(r.ra)), xb = dcross(wb, dd(r.rb));
  d3 rel = d3{vb.x + xb.x - va.x - xa.x, vb.y + xb.y - va.y - xa.y, vb.z + xb.z - va.z - xa.z};
  return d3{ddot(dd(r.t1), rel), ddot(dd(r.t2), rel), ddot(dd(r.n), rel)};
}
// |y_t| and 1 / |y_t| from s = |y_t|^2: the hardware reciprocal square root estimate refined by
// two Goldschmidt steps (g -> sqrt(s), h -> 1 / (2 sqrt(s)); each step squares the relative
// error).  s == 0 gives |y_t| = 0 (and an unused
// reciprocal); every caller leaves the sliding branch, the only user of 1 / |y_t|, when |y_t| is 0.
__device__ __forceinline__ void tangent_norm(double s, double& n, double& in) {
  const double r = __builtin_amdgcn_rsq(s);
  double g = s * r, h = 0.5 * r;
  double e = fma(-g, h, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  e = fma(-g, h, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  n = s > 0.0 ? g : 0.0;
  in = 2.0 * h;
}
// SAP projection onto the friction cone in the R norm (FP64); returns cost 1/2 gamma^T R gamma,
// writes gamma and the symmetric G = dP/dy R^-1 (00 11 22 01 02 12) in f32.  iRt = 1/Rt,
// iRn = 1/Rn.  Regions: stiction |y_t| <= mu y_n, separation y_n <= -mu_hat |y_t|, sliding.
__device__ __forceinline__ double sap_project(double mu, double Rt, double Rn, double iRt, double iRn, d3 y, d3& g,
                                              float* G) {
  double yr, iyr, yn = y.z;
  tangent_norm(y.x * y.x + y.y * y.y, yr, iyr);
  double muh = mu * Rt * iRn;
  if (yr <= mu * yn) {
    g = y;
    G[0] = (float)iRt; G[1] = G[0]; G[2] = (float)iRn; G[3] = 0.f; G[4] = 0.f; G[5] = 0.f;
  } else if (yn <= -muh * yr) {
    g = d3{0.0, 0.0, 0.0};
    G[0] = G[1] = G[2] = G[3] = G[4] = G[5] = 0.f;
  } else {
    double t0 = y.x * iyr, t1 = y.y * iyr;
    double iden = 1.0 / (1.0 + mu * muh);
    double gn = (yn + muh * yr) * iden;
    g = d3{mu * gn * t0, mu * gn * t1, gn};
    double a = mu * muh * iden, b = mu * gn * iyr;
    G[0] = (float)((a * t0 * t0 + b * (1.0 - t0 * t0)) * iRt);
    G[1] = (float)((a * t1 * t1 + b * (1.0 - t1 * t1)) * iRt);
    G[3] = (float)((a - b) * t0 * t1 * iRt);
    G[2] = (float)(iden * iRn);
    G[4] = (float)(mu * iden * t0 * iRn);
    G[5] = (float)(mu * iden * t1 * iRn);
  }
  return 0.5 * (Rt * (g.x * g.x + g.y * g.y) + Rn * g.z * g.z);
}
// Per-contact constants of the line cost below that stay fixed along a line search:
// muh = mu Rt / Rn, iden = 1 / (1 + mu muh) and the curvature coefficients a = mu muh iden,
// c2 = 2 mu iden / Rn, c3 = iden / Rn.
struct SapLineK {
  double mu, Rt, Rn, iRt, iRn;
  double muh, iden, a, c2, c3;
};
__device__ __forceinline__ SapLineK sap_line_k(double mu, double Rt, double Rn, double iRt, double iRn) {
  SapLineK k;
  k.mu = mu; k.Rt = Rt; k.Rn = Rn; k.iRt = iRt; k.iRn = iRn;
  k.muh = mu * Rt * iRn;
  k.iden = 1.0 / (1.0 + mu * k.muh);
  k.a = mu * k.muh * k.iden;
  k.c2 = 2.0 * mu * k.iden * iRn;
  k.c3 = k.iden * iRn;
  return k;
}
// Contact cost along a line: at y (velocity vc), with direction u = J dv in contact velocity,
// returns cost, adds -gamma.u to `d1` and u^T G u to `d2` (G of the FP64 projection).
__device__ __forceinline__ double sap_line(const SapLineK& k, d3 y, d3 u, double& d1, double& d2) {
  double yr, iyr, yn = y.z;
  tangent_norm(y.x * y.x + y.y * y.y, yr, iyr);
  if (yr <= k.mu * yn) {
    d1 -= y.x * u.x + y.y * u.y + y.z * u.z;
    d2 += (u.x * u.x + u.y * u.y) * k.iRt + u.z * u.z * k.iRn;
    return 0.5 * (k.Rt * (y.x * y.x + y.y * y.y) + k.Rn * y.z * y.z);
  }
  if (yn <= -k.muh * yr) return 0.0;
  double t0 = y.x * iyr, t1 = y.y * iyr;
  double gn = (yn + k.muh * yr) * k.iden;
  double mgn = k.mu * gn;
  double gt0 = mgn * t0, gt1 = mgn * t1;
  d1 -= gt0 * u.x + gt1 * u.y + gn * u.z;
  double b = mgn * iyr;
  double ut = u.x * t0 + u.y * t1;
  double uu = u.x * u.x + u.y * u.y;
  // u_t^T (a t t^T + b (I - t t^T)) u_t / Rt + 2 u_n (mu/den) t.u_t / Rn + u_n^2 / (den Rn)
  d2 += (k.a * ut * ut + b * (uu - ut * ut)) * k.iRt + k.c2 * u.z * ut + k.c3 * u.z * u.z;
  return 0.5 * (k.Rt * (gt0 * gt0 + gt1 * gt1) + k.Rn * gn * gn);
}
__device__ __forceinline__ double sap_line(double mu, double Rt, double Rn, double iRt, double iRn, d3 y, d3 u,
                                           double& d1, double& d2) {
  return sap_line(sap_line_k(mu, Rt, Rn, iRt, iRn), y, u, d1, d2);
}
// MuJoCo soft-constraint contact (contact law 1; the inner step of SolverMuJoCoAdaptive,
// mujoco_warp.step, restated from MuJoCo's computation model).  In velocity form a row with
// Jacobian c (contact frame) costs 1/2 D min(c.(w - what), 0)^2, D = 1/R; a pyramidal contact of
// condim 3 has the four rows n +- mu t_k with one common D, condim 1 (mu == 0 here) the normal row
// alone.  y = w - what.  Returns the cost, writes the impulse gamma (minus the gradient in w) and
// the Hessian G (00 11 22 01 02 12, contact frame t1 t2 n) in f32.
__device__ __forceinline__ double mj_project(double mu, double D, d3 y, d3& g, float* G) {
  g = d3{0.0, 0.0, 0.0};
  double cost = 0.0, gnn = 0.0, g00 = 0.0, g11 = 0.0, g02 = 0.0, g12 = 0.0;
  if (mu == 0.0) {
    if (y.z < 0.0) { cost = 0.5 * D * y.z * y.z; g.z = -D * y.z; gnn = D; }
  } else {
    const double e0p = y.z + mu * y.x, e0m = y.z - mu * y.x, e1p = y.z + mu * y.y, e1m = y.z - mu * y.y;
    const double Dm = D * mu, Dm2 = Dm * mu;
    if (e0p < 0.0) { cost += 0.5 * D * e0p * e0p; g.z -= D * e0p; g.x -= Dm * e0p; gnn += D; g00 += Dm2; g02 += Dm; }
    if (e0m < 0.0) { cost += 0.5 * D * e0m * e0m; g.z -= D * e0m; g.x += Dm * e0m; gnn += D; g00 += Dm2; g02 -= Dm; }
    if (e1p < 0.0) { cost += 0.5 * D * e1p * e1p; g.z -= D * e1p; g.y -= Dm * e1p; gnn += D; g11 += Dm2; g12 += Dm; }
    if (e1m < 0.0) { cost += 0.5 * D * e1m * e1m; g.z -= D * e1m; g.y += Dm * e1m; gnn += D; g11 += Dm2; g12 -= Dm; }
  }
  G[0] = (float)g00; G[1] = (float)g11; G[2] = (float)gnn; G[3] = 0.f; G[4] = (float)g02; G[5] = (float)g12;
  return cost;
}
// The same rows along a line: at y with direction u (contact velocity), returns the cost and
// adds its slope to d1 and its curvature to d2.
__device__ __forceinline__ double mj_line(double mu, double D, d3 y, d3 u, double& d1, double& d2) {
  double cost = 0.0;
  if (mu == 0.0) {
    if (y.z < 0.0) { cost = 0.5 * D * y.z * y.z; d1 += D * y.z * u.z; d2 += D * u.z * u.z; }
    return cost;
  }
  const double x[4] = {y.z + mu * y.x, y.z - mu * y.x, y.z + mu * y.y, y.z - mu * y.y};
  const double ue[4] = {u.z + mu * u.x, u.z - mu * u.x, u.z + mu * u.y, u.z - mu * u.y};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (x[e] < 0.0) {
      cost += 0.5 * D * x[e] * x[e];
      d1 += D * x[e] * ue[e];
      d2 += D * ue[e] * ue[e];
    }
  }
  return cost;
}

__device__ __forceinline__ f3 sym3_mul(const float* G, f3 u) {
  return mk(G[0] * u.x + G[3] * u.y + G[4] * u.z, G[3] * u.x + G[1] * u.y + G[5] * u.z,
            G[4] * u.x + G[5] * u.y + G[2] * u.z);
}

// ---------------------------------------------------------------- contact -> body sums
// Entries of the solve's contacts per body, built once per solve: body lane b scans the contact
// bodies (four per 32-bit LDS word, read by every lane at once) twice, counting and then listing
// the contacts that touch it in contact order; the offsets come from a wave scan.  Static sides
// belong to no lane.  Requires X.ra / X.rb (the solve's lever arms).
__device__ void build_entries(int nb, int nc, const ContactLDS& C, SolveLDS& X) {
  const int l = lane_id();
  const int me = l < nb ? l : -2;  // -1 is the static side
  const int nw = (nc + 3) >> 2;
  const int* wa = reinterpret_cast<const int*>(C.ba);
  const int* wb = reinterpret_cast<const int*>(C.bb);
  // pass 1: this body's entry count, four words (16 contacts) per step with the loads in flight
  // together (past nc the bytes are masked out)
  int deg = 0;
  for (int w0 = 0; w0 < nw; w0 += 4) {
    int a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = w0 + i < nw ? wa[w0 + i] : 0x7f7f7f7f;
      b[i] = w0 + i < nw ? wb[w0 + i] : 0x7f7f7f7f;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool in = 4 * (w0 + i) + j < nc;
        deg += (in && (int)(int8_t)((a[i] >> (8 * j)) & 0xff) == me) + (in && (int)(int8_t)((b[i] >> (8 * j)) & 0xff) == me);
      }
    }
  }
  const int off = wave_incl_scan(deg) - deg;
  // pass 2: the entries in contact order, with the side's lever arm and sign
  int k = off;
  for (int w0 = 0; w0 < nw && k < off + deg; w0 += 4) {
    int a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = w0 + i < nw ? wa[w0 + i] : 0x7f7f7f7f;
      b[i] = w0 + i < nw ? wb[w0 + i] : 0x7f7f7f7f;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = 4 * (w0 + i) + j;
        if (c < nc) {
          if ((int)(int8_t)((a[i] >> (8 * j)) & 0xff) == me) {
            X.ent[k] = (uint8_t)c;
            X.eslot[c][0] = (uint8_t)k;
            *reinterpret_cast<float4*>(X.er[k]) = make_float4(X.ra[0][c], X.ra[1][c], X.ra[2][c], -1.f);
            ++k;
          }
          if ((int)(int8_t)((b[i] >> (8 * j)) & 0xff) == me) {
            X.ent[k] = (uint8_t)c;
            X.eslot[c][1] = (uint8_t)k;
            *reinterpret_cast<float4*>(X.er[k]) = make_float4(X.rb[0][c], X.rb[1][c], X.rb[2][c], 1.f);
            ++k;
          }
        }
      }
    }
  }
  if (l < NA_MAX_BODIES) {
    X.eoff[l] = (uint16_t)off;
    X.edeg[l] = (uint16_t)deg;
  }
  wave_sync();
}

// world-frame impulse sum_k F_k lam_k of a contact-frame vector (rows t1, t2, n)
__device__ __forceinline__ d3 world_d(const CRow& r, d3 lam) {
  return d3{r.t1.x * lam.x + r.t2.x * lam.y + r.n.x * lam.z, r.t1.y * lam.x + r.t2.y * lam.y + r.n.y * lam.z,
            r.t1.z * lam.x + r.t2.z * lam.y + r.n.z * lam.z};
}

// Entry sums run four entries per step: the step's indices, lever arms and contact values are
// all requested before any is used (a scheduling barrier keeps the compiler from sinking the
// loads into the arithmetic), so a body with d entries waits for 2 ceil(d / 4) LDS round trips
// instead of 2 d.  Slots past the body's last entry re-read that entry (in bounds) and add
// nothing.
#define NA_ENT_STEP 4
#define NA_LOADS_ISSUED() __builtin_amdgcn_sched_barrier(0)

// This lane's body: jt += J_b^T gamma over its entries, from the world-frame impulses in B.xb.fd
// (sign -1 on body A, +1 on body B; torque r x f).
__device__ __forceinline__ void sum_entries_d(const BodyLDS& B, const SolveLDS& X, int eo, int ed, double* jt) {
  const int end = eo + ed;
  for (int e0 = eo; e0 < end; e0 += NA_ENT_STEP) {
    int c[NA_ENT_STEP];
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) c[i] = X.ent[min(e0 + i, end - 1)];
    float4 rr[NA_ENT_STEP];
    double fx[NA_ENT_STEP], fy[NA_ENT_STEP], fz[NA_ENT_STEP];
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) {
      rr[i] = *reinterpret_cast<const float4*>(X.er[min(e0 + i, end - 1)]);
      fx[i] = B.xb.fd[c[i]][0]; fy[i] = B.xb.fd[c[i]][1]; fz[i] = B.xb.fd[c[i]][2];
    }
    NA_LOADS_ISSUED();
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) {
      const double s = e0 + i < end ? (double)rr[i].w : 0.0;
      const d3 f = d3{s * fx[i], s * fy[i], s * fz[i]};
      const d3 t = dcross(d3{rr[i].x, rr[i].y, rr[i].z}, f);
      jt[0] += f.x; jt[1] += f.y; jt[2] += f.z;
      jt[3] += t.x; jt[4] += t.y; jt[5] += t.z;
    }
  }
}

// This lane's body: acc += the wrenches in its entry slots [eo, eo + ed) of B.xb.ew.  The slots
// are contiguous, so every load of a step is issued before the first is waited for.
__device__ __forceinline__ void sum_entry_wrenches(const BodyLDS& B, int eo, int ed, float* acc) {
  const int end = eo + ed;
  for (int e0 = eo; e0 < end; e0 += NA_ENT_STEP) {
    float4 a[NA_ENT_STEP];
    float2 b[NA_ENT_STEP];
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) {
      const int e = min(e0 + i, end - 1);
      a[i] = *reinterpret_cast<const float4*>(&B.xb.ew[e][0]);
      b[i] = *reinterpret_cast<const float2*>(&B.xb.ew[e][4]);
    }
    NA_LOADS_ISSUED();
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) {
      const bool in = e0 + i < end;
      acc[0] += in ? a[i].x : 0.f; acc[1] += in ? a[i].y : 0.f; acc[2] += in ? a[i].z : 0.f;
      acc[3] += in ? a[i].w : 0.f; acc[4] += in ? b[i].x : 0.f; acc[5] += in ? b[i].y : 0.f;
    }
  }
}

// World-frame Hessian block F^T G F of a contact (G symmetric in the contact frame, packed
// 00 11 22 01 02 12 over t1 t2 n), packed xx yy zz xy xz yz.
__device__ __forceinline__ void world_block(const CRow& r, const float* G, float* W) {
  const f3 F[3] = {r.t1, r.t2, r.n};
  const float Gf[3][3] = {{G[0], G[3], G[4]}, {G[3], G[1], G[5]}, {G[4], G[5], G[2]}};
  float xx = 0.f, yy = 0.f, zz = 0.f, xy = 0.f, xz = 0.f, yz = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const f3 h = F[0] * Gf[k][0] + F[1] * Gf[k][1] + F[2] * Gf[k][2];  // (G F)_k as a world vector
    xx += F[k].x * h.x; yy += F[k].y * h.y; zz += F[k].z * h.z;
    xy += F[k].x * h.y; xz += F[k].x * h.z; yz += F[k].y * h.z;
  }
  W[0] = xx; W[1] = yy; W[2] = zz; W[3] = xy; W[4] = xz; W[5] = yz;
}

// This lane's body: L (6x6 lower, row-major packed) += sum over its entries of
// J_b^T G J_b = [[Gw, -Gw [r]x], [[r]x Gw, -[r]x Gw [r]x]], J_b = F [I, -[r]x].
__device__ __forceinline__ void sum_entries_block(const BodyLDS& B, const SolveLDS& X, int eo, int ed, float* L) {
  const int end = eo + ed;
  for (int e0 = eo; e0 < end; e0 += NA_ENT_STEP) {
    int c[NA_ENT_STEP];
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) c[i] = X.ent[min(e0 + i, end - 1)];
    float4 rq[NA_ENT_STEP], g0q[NA_ENT_STEP], g1q[NA_ENT_STEP];
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) {
      rq[i] = *reinterpret_cast<const float4*>(X.er[min(e0 + i, end - 1)]);
      g0q[i] = *reinterpret_cast<const float4*>(&B.xb.gw[c[i]][0]);
      g1q[i] = *reinterpret_cast<const float4*>(&B.xb.gw[c[i]][4]);
    }
    NA_LOADS_ISSUED();
#pragma unroll
    for (int i = 0; i < NA_ENT_STEP; ++i) {
      const float m = e0 + i < end ? 1.f : 0.f;
      const float4 g0 = make_float4(m * g0q[i].x, m * g0q[i].y, m * g0q[i].z, m * g0q[i].w);
      const float4 g1 = make_float4(m * g1q[i].x, m * g1q[i].y, 0.f, 0.f);
      const float G[3][3] = {{g0.x, g0.w, g1.x}, {g0.w, g0.y, g1.y}, {g1.x, g1.y, g0.z}};
      const float rx = rq[i].x, ry = rq[i].y, rz = rq[i].z;
      // M = [r]x Gw, rows ([r]x row i) . columns of Gw
      float M[3][3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        M[0][j] = -rz * G[1][j] + ry * G[2][j];
        M[1][j] = rz * G[0][j] - rx * G[2][j];
        M[2][j] = -ry * G[0][j] + rx * G[1][j];
      }
      // S = -M [r]x: S[i][j] = -(M[i][0] X[0][j] + M[i][1] X[1][j] + M[i][2] X[2][j])
      const float S00 = -(M[0][1] * rz - M[0][2] * ry), S11 = -(-M[1][0] * rz + M[1][2] * rx);
      const float S22 = -(M[2][0] * ry - M[2][1] * rx);
      const float S10 = -(M[1][1] * rz - M[1][2] * ry), S20 = -(M[2][1] * rz - M[2][2] * ry);
      const float S21 = -(-M[2][0] * rz + M[2][2] * rx);
      L[0] += G[0][0]; L[1] += G[1][0]; L[2] += G[1][1]; L[3] += G[2][0]; L[4] += G[2][1]; L[5] += G[2][2];
      L[6] += M[0][0]; L[7] += M[0][1]; L[8] += M[0][2];
      L[10] += M[1][0]; L[11] += M[1][1]; L[12] += M[1][2];
      L[15] += M[2][0]; L[16] += M[2][1]; L[17] += M[2][2];
      L[9] += S00; L[13] += S10; L[14] += S11; L[18] += S20; L[19] += S21; L[20] += S22;
    }
  }
}

// 6x6 Cholesky of the packed lower triangle, in place; the diagonal slots hold the
// reciprocals 1/L_jj so that the solves only multiply.  Returns false if not SPD.
#define LI(i, j) L[(i) * ((i) + 1) / 2 + (j)]
__device__ __forceinline__ bool chol6(float* L) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    float s = LI(j, j);
#pragma unroll
    for (int k = 0; k < j; ++k) s -= LI(j, k) * LI(j, k);
    if (!(s > 0.f)) return false;
    float id = rsqrtf(s);
    LI(j, j) = id;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      float t = LI(i, j);
#pragma unroll
      for (int k = 0; k < j; ++k) t -= LI(i, k) * LI(j, k);
      LI(i, j) = t * id;
    }
  }
  return true;
}
// x = (L L^T)^-1 b with the factor of chol6
__device__ __forceinline__ void chol6_solve(const float* L, const float* b, float* x) {
  float y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    float s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= LI(i, k) * y[k];
    y[i] = s * LI(i, i);
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    float s = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s -= LI(k, i) * x[k];
    x[i] = s * LI(i, i);
  }
}
#undef LI

// Cubic Hermite minimiser on [0, h] (N&W eq. 3.59); returns a <= 0 on the fallback conditions.
__device__ __forceinline__ double cubic_min_d(double c0, double d0, double c1, double d1, double h) {
  double theta = d0 + d1 - 3.0 * (c1 - c0) / h;
  double disc = theta * theta - d0 * d1;
  if (!isfinite(disc) || disc < 0.0) return -1.0;
  double gamma = sqrt(disc);
  double denom = d1 - d0 + 2.0 * gamma;
  if (denom == 0.0) return -1.0;
  double a = h - h * (d1 + gamma - theta) / denom;
  if (!isfinite(a) || a <= 0.0 || a > h) return -1.0;
  return a;
}

// Backtracking line search with Armijo's sufficient decrease and geometric decay: the
// "armijo_decay" variant of the reference's SolverSAPAdaptive (solver_sap_adaptive.py:459, the
// default there; its solver sap_warp is absent), restated from Drake's SAP backtracking search
// (SapSolverParameters: alpha_max 1.5, rho 0.8, sufficient decrease c = 1e-4, 40 evaluations).
// Steps alpha_max rho^k are tried in turn; the search stops at the first step whose cost rises
// above the previous one's once the previous one met Armijo's condition
// l(a) - l(0) < c a l'(0) (on a convex line function that is the best step of the sequence),
// else ends with the last step if it meets the condition.  `delta(alpha)` returns the
// wave-uniform decrease l(alpha) - l(0), summed term by term (the momentum part in closed form,
// the contact costs as differences per lane) so that it is not lost in the rounding of l itself.
// Returns the step, or -1 when no evaluated step met the condition: near the optimum the
// decrease can fall below the rounding of the contact costs, and the caller then takes the exact
// search's step.
#define NA_LS_ALPHA_MAX 1.5
#define NA_LS_RHO 0.8
#define NA_LS_ARMIJO_C 1.0e-4
#define NA_LS_MAX_EVAL 40
template <class DeltaFn>
__device__ __forceinline__ double armijo_decay(double dl0, DeltaFn delta) {
  double alpha = NA_LS_ALPHA_MAX;
  double d = delta(alpha);
  bool armijo = d < NA_LS_ARMIJO_C * alpha * dl0;
  double alpha_prev = alpha, d_prev = d;
  for (int k = 1; k < NA_LS_MAX_EVAL; ++k) {
    alpha *= NA_LS_RHO;
    d = delta(alpha);
    if (d > d_prev && armijo) return alpha_prev;
    armijo = d < NA_LS_ARMIJO_C * alpha * dl0;
    alpha_prev = alpha;
    d_prev = d;
  }
  return armijo ? alpha : -1.0;
}

// exit reasons of the Newton loop
enum { NA_EXIT_CONVERGED = 0, NA_EXIT_MAX_ITER = 1, NA_EXIT_NOT_DESCENT = 2, NA_EXIT_CHOL = 3, NA_EXIT_NAN = 4 };
struct SolveStats {
  int iterations, factorizations, cg_iterations, reason;
  bool converged;
  float residual;  // |D g| / max(|D p|, |D J^T gamma|) at exit
  long long t_eval, t_pre, t_cg, t_ls;  // shader-clock cycles per phase (NA_PHASE_TIMERS builds)
};

// Phase cycle counters: compiled in only for diagnostic builds (-DNA_PHASE_TIMERS).
#ifdef NA_PHASE_TIMERS
#define NA_TIC(v) long long v = clock64()
#define NA_TOC(dst, v) (dst) += clock64() - (v)
#else
#define NA_TIC(v)
#define NA_TOC(dst, v)
#endif

// Set up the per-contact data of a step h: frame, lever arms, regularisation, v_hat.
// Requires B.x, B.iwi, B.im filled for the step's start pose.
__device__ void contact_setup(int nc, float h, const NaConfig& cfg, const ContactLDS& C, BodyLDS& B, SolveLDS& X) {
  for (int c = lane_id(); c < nc; c += NA_WAVE) {
    f3 n = mk(C.n[0][c], C.n[1][c], C.n[2][c]);
    f3 e = fabsf(n.x) < fabsf(n.y) ? (fabsf(n.x) < fabsf(n.z) ? mk(1.f, 0.f, 0.f) : mk(0.f, 0.f, 1.f))
                                   : (fabsf(n.y) < fabsf(n.z) ? mk(0.f, 1.f, 0.f) : mk(0.f, 0.f, 1.f));
    f3 t1 = nrm3(cross3(n, e));
    X.t1[0][c] = t1.x; X.t1[1][c] = t1.y; X.t1[2][c] = t1.z;
    f3 p = mk(C.p[0][c], C.p[1][c], C.p[2][c]);
    float wsum = 0.f;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      int b = s == 0 ? C.ba[c] : C.bb[c];
      f3 r = mk(0.f, 0.f, 0.f);
      if (b >= 0) {
        r = p - mk(B.x[0][b], B.x[1][b], B.x[2][b]);
        float Ii[6] = {B.iwi[0][b], B.iwi[1][b], B.iwi[2][b], B.iwi[3][b], B.iwi[4][b], B.iwi[5][b]};
        float Ir[3];
        float rr[3] = {r.x, r.y, r.z};
        sym_mul<float>(Ii, rr, Ir);
        wsum += 3.f * B.im[b] + dot3(r, r) * (Ii[0] + Ii[1] + Ii[2]) - (r.x * Ir[0] + r.y * Ir[1] + r.z * Ir[2]);
      }
      if (s == 0) { X.ra[0][c] = r.x; X.ra[1][c] = r.y; X.ra[2][c] = r.z; }
      else { X.rb[0][c] = r.x; X.rb[1][c] = r.y; X.rb[2][c] = r.z; }
    }
    float wd = wsum * (1.f / 3.f);
    float rn_soft = 1.f / (h * C.k[c] * (h + cfg.tau_d));
    float rn_rigid = cfg.beta * cfg.beta / (4.f * 9.8696044010893586f) * wd;
    float Rn = fmaxf(rn_rigid, rn_soft), Rt = cfg.sigma * wd;
    X.Rn[c] = Rn;
    X.Rt[c] = Rt;
    X.iRn[c] = 1.0 / (double)Rn;
    X.iRt[c] = 1.0 / (double)Rt;
    X.vhat[c] = -C.phi0[c] / (h + cfg.tau_d);
  }
  wave_sync();
}

// Per-contact data of a step h under the MuJoCo contact law (1): MuJoCo's contact frame (t1 from
// the y axis, or the z axis when |n_y| >= 1/2, made orthogonal to n: orthogonals,
// solvers/mujoco/kernels.py:47-58), lever arms, and for the rows of each contact
//   K, B     from solref, the time constant floored at 2h (refsafe), or the direct form;
//   R        = max(MJ_MINVAL, invweight (1 - imp)/imp) with invweight the two bodies'
//              translational inverse weights w, times (1 + mu^2) 2 mu^2 / impratio for the
//              pyramid's rows (condim 3);
//   what     = (1 - h B) J v0 - h K imp pos n, the contact velocity the rows drive towards
//              (v0: the step's start velocity, in B.v; pos: dist - margins, or the surface
//              distance where margins are zeroed).
// Rows at pos >= 0 are inactive (D = 0).  Stores D in X.Rn, what in (X.iRt, X.iRn, X.vhat).
// Requires B.x, B.im and B.v (with the static slot zero) for the step's start.
__device__ void contact_setup_mj(int nc, float h, const NaConfig& cfg, int zero_margins, const ContactLDS& C,
                                 const MjContactParams<NA_MAX_CONTACTS>& P, BodyLDS& B, SolveLDS& X) {
  for (int c = lane_id(); c < nc; c += NA_WAVE) {
    f3 n = mk(C.n[0][c], C.n[1][c], C.n[2][c]);
    f3 e = (-0.5f < n.y && n.y < 0.5f) ? mk(0.f, 1.f, 0.f) : mk(0.f, 0.f, 1.f);
    f3 t1 = nrm3(e - n * dot3(n, e));
    X.t1[0][c] = t1.x; X.t1[1][c] = t1.y; X.t1[2][c] = t1.z;
    f3 t2 = cross3(n, t1);
    f3 p = mk(C.p[0][c], C.p[1][c], C.p[2][c]);
    double w = 0.0;
    f3 rel = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      int b = s == 0 ? C.ba[c] : C.bb[c];
      f3 r = mk(0.f, 0.f, 0.f);
      if (b >= 0) {
        r = p - mk(B.x[0][b], B.x[1][b], B.x[2][b]);
        w += (double)B.im[b];
        f3 vb = mk(B.v[0][b], B.v[1][b], B.v[2][b]) + cross3(mk(B.v[3][b], B.v[4][b], B.v[5][b]), r);
        rel = s == 0 ? rel - vb : rel + vb;
      }
      if (s == 0) { X.ra[0][c] = r.x; X.ra[1][c] = r.y; X.ra[2][c] = r.z; }
      else { X.rb[0][c] = r.x; X.rb[1][c] = r.y; X.rb[2][c] = r.z; }
    }
    const double pos = zero_margins ? (double)C.dist[c] : (double)C.phi0[c];
    const double imp = (double)C.k[c], mu = (double)C.mu[c], dmax = (double)P.dmax[c];
    const double tc0 = (double)P.tc[c], dr = (double)P.dr[c], hd = (double)h;
    double K, Bd;
    if (tc0 > 0.0) {
      const double tc = fmax(tc0, 2.0 * hd);
      K = 1.0 / (dmax * dmax * tc * tc * dr * dr);
      Bd = 2.0 / (dmax * tc);
    } else {
      K = -tc0 / (dmax * dmax);
      Bd = -dr / dmax;
    }
    const double wr = mu > 0.0 ? w * (1.0 + mu * mu) * 2.0 * mu * mu / (double)cfg.impratio : w;
    const double R = fmax((double)NA_MJ_MINVAL, wr * (1.0 - imp) / imp);
    const double sv = 1.0 - hd * Bd;
    X.Rn[c] = pos < 0.0 ? (float)(1.0 / R) : 0.f;
    X.Rt[c] = 0.f;
    X.iRt[c] = sv * (double)dot3(t1, rel);
    X.iRn[c] = sv * (double)dot3(t2, rel);
    X.vhat[c] = (float)(sv * (double)dot3(n, rel) - hd * K * imp * pos);
  }
  wave_sync();
}

// Contact passes.  Lane l handles contacts l and l + 64 (at most NA_MAX_CONTACTS = 128); the
// NG = 1 or 2 groups are loaded together before any arithmetic, so one wait covers all
// gathers.  Values a pass keeps in registers for later passes follow the same assignment.
static_assert(NA_MAX_CONTACTS <= 2 * NA_WAVE, "contact passes assume at most two contacts per lane");

// One contact's constants for a solve (regularisation, its FP64 reciprocals, stabilisation
// velocity, friction), loaded once into registers.  Under the MuJoCo law (1) the same slots hold
// the target contact velocity (iRt, iRn, vh = t1, t2, n components) and the rows' D (Rn).
struct CConst {
  double iRt, iRn;
  float Rt, Rn, vh, mu;
};
__device__ __forceinline__ CConst load_const(const ContactLDS& C, const SolveLDS& X, int c) {
  return CConst{X.iRt[c], X.iRn[c], X.Rt[c], X.Rn[c], X.vhat[c], C.mu[c]};
}

// FP64 evaluation at the iterate in B.vd (visible to all lanes) for this lane's contact rows
// `r`: the f32 blocks G into `G` and, when `gather` is set, the contact velocities into `vc`
// (otherwise `vc` already holds them), both in registers; this lane's body gets J^T gamma in
// `jt` (entries [eo, eo + ed)).  Returns this lane's part of the contact cost.
template <int NG, int LAW>
__device__ __forceinline__ double eval_contacts_g(int nc, BodyLDS& B, const SolveLDS& X, const CRow* r,
                                                  const CConst* k, float (*G)[6], d3* vc, bool gather, int eo,
                                                  int ed, double* jt) {
  const int l = lane_id();
  if (gather) {
#pragma unroll
    for (int g = 0; g < NG; ++g) vc[g] = row_vel_d(B, r[g]);
  }
  double cost = 0.0;
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int c = l + g * NA_WAVE;
    const bool valid = c < nc;
    const double iRt = k[g].iRt, iRn = k[g].iRn;
    d3 gam;
    double cc;
    if (LAW == 1) {
      d3 y = d3{vc[g].x - iRt, vc[g].y - iRn, vc[g].z - (double)k[g].vh};
      cc = mj_project(k[g].mu, k[g].Rn, y, gam, G[g]);
    } else {
      d3 y = d3{-vc[g].x * iRt, -vc[g].y * iRt, -(vc[g].z - (double)k[g].vh) * iRn};
      cc = sap_project(k[g].mu, k[g].Rt, k[g].Rn, iRt, iRn, y, gam, G[g]);
    }
    cost += valid ? cc : 0.0;
    if (valid) {
      const d3 fw = world_d(r[g], gam);
      B.xb.fd[c][0] = fw.x; B.xb.fd[c][1] = fw.y; B.xb.fd[c][2] = fw.z;
    }
  }
  wave_sync();
#pragma unroll
  for (int q = 0; q < 6; ++q) jt[q] = 0.0;
  sum_entries_d(B, X, eo, ed, jt);
  return cost;
}

// Block-Jacobi blocks: this lane's body gets L += sum over its contacts of J_b^T G J_b (the
// world-frame blocks F^T G F of every contact stored in B.xb.gw first).  The caller has made
// every lane's reads of B.xb (J^T gamma) complete.
template <int NG>
__device__ __forceinline__ void precond_blocks_g(int nc, BodyLDS& B, const SolveLDS& X, const CRow* r,
                                                 const float (*G)[6], int eo, int ed, float* L) {
  const int l = lane_id();
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int c = l + g * NA_WAVE;
    if (c < nc) {
      float W[6];
      world_block(r[g], G[g], W);
      *reinterpret_cast<float4*>(&B.xb.gw[c][0]) = make_float4(W[0], W[1], W[2], W[3]);
      *reinterpret_cast<float4*>(&B.xb.gw[c][4]) = make_float4(W[4], W[5], 0.f, 0.f);
    }
  }
  wave_sync();
  sum_entries_block(B, X, eo, ed, L);
}

// One contact's line-search constants, held in registers for every trial step of a Newton
// iteration: start velocity w0 and direction dv in the contact frame, the regularisation and
// its reciprocals, friction and the stabilisation velocity.
struct LsRow {
  d3 w0, dv;
  double iRt, iRn, Rt, Rn, mu, vh;
  SapLineK k;  // contact law 0
};

// Line-search setup: contact-space direction dv = J dx (dx in B.vd) into this lane's rows `lr`,
// and the lane's contact terms at alpha = 0 (cost lc0, slope d1c0, curvature d2c0) and alpha = 1
// (cost lc1, slope d1c1).
template <int NG, int LAW>
__device__ __forceinline__ void ls_setup_g(int nc, const BodyLDS& B, const CRow* r, const CConst* k, const d3* vc0,
                                           LsRow* lr, double& lc0, double& d1c0, double& d2c0, double& lc1,
                                           double& d1c1) {
  const int l = lane_id();
#pragma unroll
  for (int g = 0; g < NG; ++g) lr[g].dv = row_vel_d(B, r[g]);
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    LsRow& q = lr[g];
    q.iRt = k[g].iRt; q.iRn = k[g].iRn; q.Rt = k[g].Rt; q.Rn = k[g].Rn; q.mu = k[g].mu;
    q.vh = (double)k[g].vh;
    q.w0 = vc0[g];
    if (LAW == 0) q.k = sap_line_k(q.mu, q.Rt, q.Rn, q.iRt, q.iRn);
  }
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int c = l + g * NA_WAVE;
    const LsRow& q = lr[g];
    const d3 w0 = q.w0, u = q.dv;
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0, ca, cb;
    if (LAW == 1) {
      d3 y0 = d3{w0.x - q.iRt, w0.y - q.iRn, w0.z - q.vh};
      d3 y1 = d3{y0.x + u.x, y0.y + u.y, y0.z + u.z};
      ca = mj_line(q.mu, q.Rn, y0, u, a1, a2);
      cb = mj_line(q.mu, q.Rn, y1, u, b1, b2);
    } else {
      d3 y0 = d3{-w0.x * q.iRt, -w0.y * q.iRt, -(w0.z - q.vh) * q.iRn};
      d3 y1 = d3{-(w0.x + u.x) * q.iRt, -(w0.y + u.y) * q.iRt, -(w0.z + u.z - q.vh) * q.iRn};
      ca = sap_line(q.k, y0, u, a1, a2);
      cb = sap_line(q.k, y1, u, b1, b2);
    }
    if (c < nc) {
      lc0 += ca;
      d1c0 += a1;
      d2c0 += a2;
      lc1 += cb;
      d1c1 += b1;
    }
  }
}

// Contact part of the line function at alpha for this lane's contacts (FP64), from the rows
// that ls_setup_g left in registers: adds the cost, its first and second derivatives.
template <int NG, int LAW>
__device__ __forceinline__ void ls_contacts_g(int nc, const LsRow* lr, double alpha, double& l, double& dl,
                                              double& d2l) {
  const int ln = lane_id();
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int c = ln + g * NA_WAVE;
    const LsRow& q = lr[g];
    const d3 dv = q.dv;
    double a = 0.0, b = 0.0, cc;
    if (LAW == 1) {
      d3 y = d3{q.w0.x + alpha * dv.x - q.iRt, q.w0.y + alpha * dv.y - q.iRn, q.w0.z + alpha * dv.z - q.vh};
      cc = mj_line(q.mu, q.Rn, y, dv, a, b);
    } else {
      d3 y = d3{-(q.w0.x + alpha * dv.x) * q.iRt, -(q.w0.y + alpha * dv.y) * q.iRt,
                -(q.w0.z + alpha * dv.z - q.vh) * q.iRn};
      cc = sap_line(q.k, y, dv, a, b);
    }
    if (c < nc) {
      l += cc;
      dl += a;
      d2l += b;
    }
  }
}

// Preconditioned CG on H x = -g * gscale (f32) with the block-Jacobi factor L of this lane's
// body; stops when the preconditioned residual r.z falls to eta2 (the squared relative
// tolerance) times its start value, or after max_cg iterations.  The lane's
// contact rows `rw` and G blocks `G` are the caller's registers.  The product H p = A p + J^T G J p:
// each contact lane writes the wrench of its world-frame force f = F^T G J p on each dynamic side
// into that side's entry slot, (-f, -(r_A x f)) and (f, r_B x f), and each body lane sums its
// slots [eo, eo + ed).  Returns the iteration count.
template <int NG>
__device__ __forceinline__ int pcg(BodyLDS& B, const SolveLDS& X, int nc, const CRow* rw, const float (*G)[6],
                                   const BodyReg& br, bool own, int eo, int ed, const float* L, const double* g,
                                   double gscale, float eta2, int max_cg, float* x) {
  const int l = lane_id();
  float r[6], z[6], pp[6];
  float rz = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) { x[k] = 0.f; r[k] = (float)(-g[k] * gscale); z[k] = 0.f; }
  if (own) chol6_solve(L, r, z);
#pragma unroll
  for (int k = 0; k < 6; ++k) { pp[k] = z[k]; rz += r[k] * z[k]; }
  rz = wave_sum(rz);
  const float tol2 = eta2 * rz;
  int it = 0;
  for (; it < max_cg && rz > tol2 && rz > 0.f; ++it) {
    wave_sync();  // the body lanes' reads of the last product's entries precede this product's stores
    f3 u[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) u[q] = row_vel_lanes(pp, rw[q]);
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      const int c = l + q * NA_WAVE;
      if (c < nc) {
        const f3 s3 = sym3_mul(G[q], u[q]);
        const f3 fw = rw[q].t1 * s3.x + rw[q].t2 * s3.y + rw[q].n * s3.z;
        if (rw[q].ha) {
          const f3 fa = mk(-fw.x, -fw.y, -fw.z);
          const f3 ta = cross3(rw[q].ra, fa);
          *reinterpret_cast<float4*>(&B.xb.ew[rw[q].ea][0]) = make_float4(fa.x, fa.y, fa.z, ta.x);
          *reinterpret_cast<float2*>(&B.xb.ew[rw[q].ea][4]) = make_float2(ta.y, ta.z);
        }
        if (rw[q].hb) {
          const f3 tb = cross3(rw[q].rb, fw);
          *reinterpret_cast<float4*>(&B.xb.ew[rw[q].eb][0]) = make_float4(fw.x, fw.y, fw.z, tb.x);
          *reinterpret_cast<float2*>(&B.xb.ew[rw[q].eb][4]) = make_float2(tb.y, tb.z);
        }
      }
    }
    wave_sync();
    float Hp[6];
    float pHp = 0.f;
    if (own) {
      float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      sum_entry_wrenches(B, eo, ed, acc);
      applyA<float>(br, pp, Hp);
#pragma unroll
      for (int q = 0; q < 6; ++q) { Hp[q] += acc[q]; pHp += pp[q] * Hp[q]; }
    } else {
#pragma unroll
      for (int q = 0; q < 6; ++q) Hp[q] = 0.f;
    }
    pHp = wave_sum(pHp);
    if (!(pHp > 0.f)) break;
    float alpha = rz * __builtin_amdgcn_rcpf(pHp);
    float rzn = 0.f;
#pragma unroll
    for (int q = 0; q < 6; ++q) { x[q] += alpha * pp[q]; r[q] -= alpha * Hp[q]; }
    if (own) chol6_solve(L, r, z);
#pragma unroll
    for (int q = 0; q < 6; ++q) rzn += r[q] * z[q];
    rzn = wave_sum(rzn);
    float beta = rzn * __builtin_amdgcn_rcpf(rz);
    rz = rzn;
#pragma unroll
    for (int q = 0; q < 6; ++q) pp[q] = z[q] + beta * pp[q];
  }
  return it;
}

// Newton solve with FP64 residuals and FP32 PCG directions.  Lane b < nb owns body b:
// `br` its constants, `v` its velocity (in: guess, out: solution).  Returns true when
// the SAP optimality criterion holds.  With cfg.cg_forcing > 0 the PCG tolerance of Newton
// iteration k is eta_k = min(cg_forcing, sqrt(|Dg_k| / |Dg_0|)), never tighter than needed
// to reach the optimality target in one step (0.1 target / |Dg_k|) nor than cg_rel_tol.
// NG (1 or 2) is the number of contacts per lane; each lane's contact rows, G blocks and
// contact velocities stay in registers for the whole solve.
template <int NG, int LAW>
__device__ __forceinline__ bool sap_solve_g(int nb, int nc, const NaConfig& cfg, const ContactLDS& C, BodyLDS& B,
                                            const SolveLDS& X, const BodyReg& br, double* v, SolveStats& st) {
  nb = uni_i(nb);
  nc = uni_i(nc);
  const int l = lane_id();
  const bool own = l < nb;
  st.iterations = 0;
  st.factorizations = 0;
  st.cg_iterations = 0;
  st.converged = false;
  st.reason = NA_EXIT_MAX_ITER;
  st.residual = 0.f;
  st.t_eval = st.t_pre = st.t_cg = st.t_ls = 0;
  // settings in scalar registers for the loops below
  const int max_iterations = uni_i(cfg.max_iterations), max_cg = uni_i(cfg.max_cg);
  const double opt_abs = uni_f(cfg.opt_abs_tol), opt_rel = uni_f(cfg.opt_rel_tol), ls_rel = uni_f(cfg.ls_rel_tol);
  const double cg_rel = uni_f(cfg.cg_rel_tol), forcing = uni_f(cfg.cg_forcing);
  const int line_search = uni_i(cfg.line_search);
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) B.v[k][NA_DUMMY] = 0.f;
  }
  CRow rows[NG];
  float Gr[NG][6];
  d3 vc0[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) rows[g] = load_row(C, X, l + g * NA_WAVE, l + g * NA_WAVE < nc);
  bool gather = true;  // the first evaluation gathers J v from the guess
  // this lane's body entries (build_entries)
  const int eo = own ? (int)X.eoff[l] : 0, ed = own ? (int)X.edeg[l] : 0;
  CConst kc[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) kc[g] = load_const(C, X, l + g * NA_WAVE);
  double gn_first = 0.0;
  for (int it = 0; it <= max_iterations; ++it) {
    // gradient g = A (v - v*) - J^T gamma, FP64
    NA_TIC(t0);
    if (own && gather) {
#pragma unroll
      for (int k = 0; k < 6; ++k) B.vd[k][l] = v[k];
    }
    if (l == 0 && gather) {  // the static side's zero velocity
#pragma unroll
      for (int k = 0; k < 6; ++k) B.vd[k][NA_DUMMY] = 0.0;
    }
    wave_sync();
    double jtg[6];
    double ccost = eval_contacts_g<NG, LAW>(nc, B, X, rows, kc, Gr, vc0, gather, eo, ed, jtg);
    double g[6], d[6], Ad[6];
    double mom = 0.0, gn = 0.0, pn = 0.0, jn = 0.0;
    if (own) {
      double p[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) d[k] = v[k] - br.vs[k];
      applyA<double>(br, d, Ad);
      applyA<double>(br, v, p);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double jc = jtg[k];
        g[k] = Ad[k] - jc;
        mom += 0.5 * d[k] * Ad[k];
        double s = br.D[k];
        gn += (s * g[k]) * (s * g[k]);
        pn += (s * p[k]) * (s * p[k]);
        jn += (s * jc) * (s * jc);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) { g[k] = 0.0; d[k] = 0.0; Ad[k] = 0.0; }
    }
    gn = sqrt(wave_sum_d(gn));
    pn = sqrt(wave_sum_d(pn));
    jn = sqrt(wave_sum_d(jn));
    const double mom0 = wave_sum_d(mom);
    const double l0 = mom0 + wave_sum_d(ccost);
    NA_TOC(st.t_eval, t0);
    st.residual = (float)(gn / fmax(fmax(pn, jn), 1e-300));
    if (!(gn == gn)) {
      st.reason = NA_EXIT_NAN;
      return false;
    }
    const double target = opt_abs + opt_rel * fmax(pn, jn);
    if (gn <= target) {
      st.converged = true;
      st.reason = NA_EXIT_CONVERGED;
      return true;
    }
    if (it == max_iterations) break;
    if (it == 0) gn_first = gn;

    // block-Jacobi preconditioner P_b = A_b + sum J_b^T G J_b (f32); the blocks reuse the
    // exchange buffer of J^T gamma, so every lane must have read its entries first
    NA_TIC(t1);
    wave_sync();
    float L[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) L[k] = 0.f;
    precond_blocks_g<NG>(nc, B, X, rows, Gr, eo, ed, L);
    bool spd = true;
    if (own) {
      L[0] += br.m; L[2] += br.m; L[5] += br.m;
      // packed indices of rows/cols 3..5: (3,3)=9 (4,3)=13 (4,4)=14 (5,3)=18 (5,4)=19 (5,5)=20
      L[9] += br.I[0]; L[14] += br.I[1]; L[20] += br.I[2];
      L[13] += br.I[3]; L[18] += br.I[4]; L[19] += br.I[5];
      spd = chol6(L);
    }
    if (__any(!spd)) {
      st.reason = NA_EXIT_CHOL;
      return false;
    }
    st.factorizations++;
    NA_TOC(st.t_pre, t1);

    // PCG on H x = -g in f32, right-hand side scaled to O(1)
    NA_TIC(t2);
    double eta = cg_rel;
    if (forcing > 0.0) {
      double e = fmin(sqrt(gn / gn_first), forcing);
      e = fmax(e, fmax(eta, 0.1 * target / gn));
      eta = fmin(e, fmax(forcing, eta));
    }
    const double gscale = gn > 0.0 ? 1.0 / gn : 1.0;
    float x[6];
    const float eta2 = (float)(eta * eta);
    st.cg_iterations += pcg<NG>(B, X, nc, rows, Gr, br, own, eo, ed, L, g, gscale, eta2, max_cg, x);
    double dx[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) dx[q] = own ? (double)x[q] / gscale : 0.0;
    NA_TOC(st.t_cg, t2);
    NA_TIC(t3);

    // exact line search along dx in FP64:  l(a) = mom0 + a b0 + a^2 aa / 2 + contacts(a).
    // One pass computes the contact-space direction and the contact terms at a = 0 and 1.
    if (own) {
#pragma unroll
      for (int q = 0; q < 6; ++q) B.vd[q][l] = dx[q];
    }
    if (l == 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) B.vd[q][NA_DUMMY] = 0.0;
    }
    wave_sync();
    double lc0 = 0.0, d1c0 = 0.0, d2c0 = 0.0, lc1 = 0.0, d1c1 = 0.0;
    LsRow lr[NG];
    ls_setup_g<NG, LAW>(nc, B, rows, kc, vc0, lr, lc0, d1c0, d2c0, lc1, d1c1);
    double aa = 0.0, b0 = 0.0;
    if (own) {
      double Ax[6];
      applyA<double>(br, dx, Ax);
#pragma unroll
      for (int q = 0; q < 6; ++q) { aa += dx[q] * Ax[q]; b0 += dx[q] * Ad[q]; }
    }
    aa = wave_sum_d(aa);
    b0 = wave_sum_d(b0);
    const double dl0 = b0 + wave_sum_d(d1c0);
    const double d2l0 = aa + wave_sum_d(d2c0);
    const double l1 = mom0 + b0 + 0.5 * aa + wave_sum_d(lc1);
    const double dl1 = b0 + aa + wave_sum_d(d1c1);
    if (!(dl0 < 0.0)) {
      st.iterations = it + 1;
      st.reason = NA_EXIT_NOT_DESCENT;
      break;
    }
    double alpha = -1.0;
    if (line_search == 1) {
      alpha = armijo_decay(dl0, [&](double a) {
        double lc = 0.0, dlc = 0.0, d2c = 0.0;
        ls_contacts_g<NG, LAW>(nc, lr, a, lc, dlc, d2c);
        return a * b0 + 0.5 * a * a * aa + wave_sum_d(lc - lc0);
      });
    }
    if (alpha > 0.0) {
      // backtracking step taken
    } else if (dl1 <= 0.0) {
      alpha = 1.0;
    } else {
      double lo = 0.0, hi = 1.0;
      alpha = cubic_min_d(l0, dl0, l1, dl1, 1.0);
      if (!(alpha > 0.0)) alpha = fmin(fmax(-dl0 / d2l0, 0.0), 1.0);
      if (!(alpha > lo && alpha < hi)) alpha = 0.5;
      for (int k = 0; k < 60; ++k) {
        double lc = 0.0, dlc = 0.0, d2c = 0.0;
        ls_contacts_g<NG, LAW>(nc, lr, alpha, lc, dlc, d2c);
        double dl = b0 + alpha * aa + wave_sum_d(dlc);
        double ddl = aa + wave_sum_d(d2c);
        if (fabs(dl) <= ls_rel * fabs(dl0)) break;
        if (dl < 0.0) lo = alpha;
        else hi = alpha;
        double an = ddl > 0.0 ? alpha - dl / ddl : -1.0;
        if (!(an > lo && an < hi)) an = 0.5 * (lo + hi);
        if (hi - lo <= 1e-14 * hi) { alpha = an; break; }
        alpha = an;
      }
    }
    if (own) {
#pragma unroll
      for (int q = 0; q < 6; ++q) v[q] += alpha * dx[q];
    }
    // the contact velocities are linear in v: the next iterate's are vc0 + alpha J dx
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      vc0[g].x += alpha * lr[g].dv.x;
      vc0[g].y += alpha * lr[g].dv.y;
      vc0[g].z += alpha * lr[g].dv.z;
    }
    gather = false;
    NA_TOC(st.t_ls, t3);
    st.iterations = it + 1;
  }
  return st.converged;
}

__device__ bool sap_solve(int nb, int nc, const NaConfig& cfg, const ContactLDS& C, BodyLDS& B, const SolveLDS& X,
                          const BodyReg& br, double* v, SolveStats& st) {
  if (cfg.contact_law == 1)
    return nc > NA_WAVE ? sap_solve_g<2, 1>(nb, nc, cfg, C, B, X, br, v, st)
                        : sap_solve_g<1, 1>(nb, nc, cfg, C, B, X, br, v, st);
  return nc > NA_WAVE ? sap_solve_g<2, 0>(nb, nc, cfg, C, B, X, br, v, st)
                      : sap_solve_g<1, 0>(nb, nc, cfg, C, B, X, br, v, st);
}

// ---------------------------------------------------------------- body setup and integration
// Per-lane constants of body (lane < nb) at pose q (f64), velocity v0, wrench f, step h.
__device__ void body_setup(const NaModel& M, int gb, const double* q, const float* v0, const float* f, double h,
                           BodyReg& br, BodyLDS& B, int l) {
  m3 R = qmat((float)q[3], (float)q[4], (float)q[5], (float)q[6]);
  const float* Ib = M.body_inertia + 9 * gb;
  const float* Ibi = M.body_inv_inertia + 9 * gb;
  float Iw[9], Iwi[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float a = 0.f, ai = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          a += R.m[3 * i + k] * Ib[3 * k + m] * R.m[3 * j + m];
          ai += R.m[3 * i + k] * Ibi[3 * k + m] * R.m[3 * j + m];
        }
      }
      Iw[3 * i + j] = a;
      Iwi[3 * i + j] = ai;
    }
  }
  br.m = M.body_mass[gb];
  br.im = M.body_inv_mass[gb];
  br.I[0] = Iw[0]; br.I[1] = Iw[4]; br.I[2] = Iw[8]; br.I[3] = Iw[1]; br.I[4] = Iw[2]; br.I[5] = Iw[5];
  float Ii[6] = {Iwi[0], Iwi[4], Iwi[8], Iwi[1], Iwi[2], Iwi[5]};
  f3 com = mk(M.body_com[3 * gb], M.body_com[3 * gb + 1], M.body_com[3 * gb + 2]);
  f3 xc = mk((float)q[0], (float)q[1], (float)q[2]) + mv(R, com);
  B.x[0][l] = xc.x; B.x[1][l] = xc.y; B.x[2][l] = xc.z;
#pragma unroll
  for (int k = 0; k < 6; ++k) B.iwi[k][l] = Ii[k];
  B.im[l] = br.im;
  double gmask = br.im > 0.f ? 1.0 : 0.0;
  double hd = h;
  br.vs[0] = (double)v0[0] + hd * ((double)f[0] * br.im + (double)M.gx * gmask);
  br.vs[1] = (double)v0[1] + hd * ((double)f[1] * br.im + (double)M.gy * gmask);
  br.vs[2] = (double)v0[2] + hd * ((double)f[2] * br.im + (double)M.gz * gmask);
  double w0[3] = {v0[3], v0[4], v0[5]}, Iw0[3];
  sym_mul<double>(br.I, w0, Iw0);
  d3 tau = d3{(double)f[3] - (w0[1] * Iw0[2] - w0[2] * Iw0[1]), (double)f[4] - (w0[2] * Iw0[0] - w0[0] * Iw0[2]),
              (double)f[5] - (w0[0] * Iw0[1] - w0[1] * Iw0[0])};
  double t3[3] = {tau.x, tau.y, tau.z}, dw[3];
  sym_mul<double>(Ii, t3, dw);
  br.vs[3] = w0[0] + hd * dw[0];
  br.vs[4] = w0[1] + hd * dw[1];
  br.vs[5] = w0[2] + hd * dw[2];
  br.D[0] = br.D[1] = br.D[2] = rsqrtf(br.m);
  br.D[3] = rsqrtf(br.I[0]);
  br.D[4] = rsqrtf(br.I[1]);
  br.D[5] = rsqrtf(br.I[2]);
}

// Semi-implicit update in double (solver.py:50-94, no damping): x_com += h v,
// q = normalize(q + 1/2 h (w, 0) q), origin = x_com - R com.  mj_quat: the rotation as MuJoCo
// integrates a free joint (mju_quatIntegrate), by the exact exponential q = exp(h w / 2) q of the
// new angular velocity, normalised.
__device__ void integrate_body(const NaModel& M, int gb, const double* q, const double* v, double h, double* qo,
                               bool mj_quat = false) {
  double qx = q[3], qy = q[4], qz = q[5], qw = q[6];
  double cx = M.body_com[3 * gb], cy = M.body_com[3 * gb + 1], cz = M.body_com[3 * gb + 2];
  double xc[3];
  xc[0] = q[0] + (1 - 2 * (qy * qy + qz * qz)) * cx + 2 * (qx * qy - qz * qw) * cy + 2 * (qx * qz + qy * qw) * cz;
  xc[1] = q[1] + 2 * (qx * qy + qz * qw) * cx + (1 - 2 * (qx * qx + qz * qz)) * cy + 2 * (qy * qz - qx * qw) * cz;
  xc[2] = q[2] + 2 * (qx * qz - qy * qw) * cx + 2 * (qy * qz + qx * qw) * cy + (1 - 2 * (qx * qx + qy * qy)) * cz;
  xc[0] += h * v[0];
  xc[1] += h * v[1];
  xc[2] += h * v[2];
  double wx = v[3], wy = v[4], wz = v[5];
  double nx, ny, nz, nw;
  if (mj_quat) {
    double wn = sqrt(wx * wx + wy * wy + wz * wz), th = 0.5 * h * wn;
    double sc = wn > 0.0 ? sin(th) / wn : 0.0, cw = cos(th);
    double ex = wx * sc, ey = wy * sc, ez = wz * sc;  // exp(h w / 2) = (e, cw), times q
    nx = cw * qx + ex * qw + ey * qz - ez * qy;
    ny = cw * qy + ey * qw + ez * qx - ex * qz;
    nz = cw * qz + ez * qw + ex * qy - ey * qx;
    nw = cw * qw - ex * qx - ey * qy - ez * qz;
  } else {
    double dx = wy * qz - wz * qy + wx * qw;
    double dy = wz * qx - wx * qz + wy * qw;
    double dz = wx * qy - wy * qx + wz * qw;
    double dw = -(wx * qx + wy * qy + wz * qz);
    nx = qx + 0.5 * h * dx; ny = qy + 0.5 * h * dy; nz = qz + 0.5 * h * dz; nw = qw + 0.5 * h * dw;
  }
  double il = 1.0 / sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
  nx *= il; ny *= il; nz *= il; nw *= il;
  qo[0] = xc[0] - ((1 - 2 * (ny * ny + nz * nz)) * cx + 2 * (nx * ny - nz * nw) * cy + 2 * (nx * nz + ny * nw) * cz);
  qo[1] = xc[1] - (2 * (nx * ny + nz * nw) * cx + (1 - 2 * (nx * nx + nz * nz)) * cy + 2 * (ny * nz - nx * nw) * cz);
  qo[2] = xc[2] - (2 * (nx * nz - ny * nw) * cx + 2 * (ny * nz + nx * nw) * cy + (1 - 2 * (nx * nx + ny * ny)) * cz);
  qo[3] = nx; qo[4] = ny; qo[5] = nz; qo[6] = nw;
}
