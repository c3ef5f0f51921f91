// Articulated worlds inside one wavefront: the generalized-coordinate step of a tree of revolute
// joints under an optional FREE root (floating base), with contacts.
//
// Coordinates follow Newton's joint_q / joint_qd (state.py:149-153): a REVOLUTE joint has one
// angle / rate; the FREE root has (origin position, quaternion xyzw) / (COM velocity, angular
// velocity), world frame.  Joint j moves body j and parents come before children.  Spatial
// vectors are world-frame Pluecker coordinates at the world origin, (angular; linear).
// For a step of size h from (q, v):
//   M(q)     composite-rigid-body mass matrix, FP64, one lane per dof column
//   c(q, v)  Newton-Euler velocity products and gravity, FP64
//   A = M + diag(armature + h (kd + h ke)),
//   A v* = (M + diag(armature)) v + h (ke (q_t - q) + kd qd_t + tau - c)
// (drive targets q_t, qd_t and applied generalized forces tau from the control; applied body
// wrenches enter the bias c with gravity)
//   H = A + sum_c J_c^T G_c J_c   with the contact part on FP32 MFMA (v_mfma_f32_32x32x2f32)
// and the convex contact problem of the free-body solve with this dense A: Newton steps on the
// FP64 gradient, Cholesky factors in LDS, FP64 substitutions, the exact line search of
// na_solve.h.  Contact rows are J_c = F^T (P_b J_b - P_a J_a), P the point-velocity map at the
// contact point; R_n and R_t use w = tr(J_c A^-1 J_c^T) / 3, and a contact with w below 1e-10
// 1/kg (a point that no joint can move, e.g. on the axis of a fixed-base joint) imposes nothing.
// Layout: lane k owns dof k (velocity, gradient), lane b body b (kinematics), lane c contact c.
#pragma once
#include "na_solve.h"

#define NA_ART_NB 16            // bodies (= joints) per articulated world
#define NA_ART_ND 32            // dofs per world (one MFMA tile)
#define NA_ART_NQ 32            // coordinates per world
#define NA_ART_NC 32            // contacts per world
#define NA_ART_LD 33            // row stride of the 32 x 32 LDS matrices
#define NA_ART_CC 16            // contacts per MFMA chunk: 48 reduction rows
#define NA_ART_DUMMY NA_ART_NB  // body slot of the static side: zero velocity, no dofs
#define NA_JT_REVOLUTE 1
#define NA_JT_FREE 4
#define NA_ART_W_MIN 1.0e-10

typedef float na_v16f __attribute__((ext_vector_type(16)));

struct ArtLDS {
  // contact generation, the MFMA operands of the contact Hessian and the FP64 factor of A are
  // live in disjoint phases of a step
  union {
    SegLDS seg;
    struct {
      float P[3 * NA_ART_CC][32];  // contact rows J_c (row 3c + r, column = dof)
      float Q[3 * NA_ART_CC][32];  // G_c J_c
    } mm;
    struct {
      double LA[NA_ART_ND][NA_ART_LD];  // Cholesky factor of A (lower)
      double wrow[3 * NA_ART_NC];       // |L^-1 J_r^T|^2 per contact row (Delassus diagonal)
    } fa;
  };
  ContactSet<NA_ART_NC> con;
  // joint data of the world
  int jtype[NA_ART_NB], jpar[NA_ART_NB], jq[NA_ART_NB], jd[NA_ART_NB], depth[NA_ART_NB];
  uint32_t anc[NA_ART_NB + 1];  // bit a set: body a is this body or one of its ancestors
  int dof_body[NA_ART_ND];
  int dof_q[NA_ART_ND];         // coordinate of a revolute dof, -1 for FREE dofs
  float Xp[NA_ART_NB][7], Xc[NA_ART_NB][7], axis[NA_ART_NB][3], com_b[NA_ART_NB][3];
  float kp[NA_ART_ND], kd[NA_ART_ND], arm[NA_ART_ND], qt[NA_ART_ND];
  float vt[NA_ART_ND];          // velocity target of the drive (0 without control)
  float tau[NA_ART_ND];         // applied generalized force: revolute joint_f + joint_act
  float fext[NA_ART_NB][6];     // applied wrench at the body's COM, world frame (f; t): body_f + FREE joint_f
  // kinematics at the evaluation coordinates qe
  double qe[NA_ART_NQ];
  double Rb[NA_ART_NB][9], pb[NA_ART_NB][3];  // body pose (rotation row-major, origin)
  double cb[NA_ART_NB][3], Ib[NA_ART_NB][6];  // COM and world inertia about it (xx yy zz xy xz yz)
  double mb[NA_ART_NB];                       // mass; during the mass matrix: the subtree's
  float bqf[7][NA_ART_NB];                    // f32 poses for contact generation
  double S[NA_ART_ND][6];                     // motion subspace columns
  double V[NA_ART_NB + 1][6];                 // body velocities (body forces inside the bias)
  double T[NA_ART_ND][6];                     // per-dof velocity products, then composite inertia x S
  // matrices
  double A[NA_ART_ND][NA_ART_LD];  // M, then M + diag(armature + h (kd + h ke))
  float H[NA_ART_ND][NA_ART_LD];   // Newton Hessian and its Cholesky factor
  double x[NA_ART_ND];             // FP64 broadcast vector
  double qo[NA_ART_NQ];            // integrated coordinates
  // contacts of the step
  float t1[3][NA_ART_NC];
  double Rn[NA_ART_NC], Rt[NA_ART_NC], iRn[NA_ART_NC], iRt[NA_ART_NC], vhat[NA_ART_NC];
  double vc0[3][NA_ART_NC], dvc[3][NA_ART_NC];
  double W[6][NA_ART_NC];  // impulse (p x f; f) of the contact force at the world origin
  float G[6][NA_ART_NC];
  int act[NA_ART_NC];
};

// Control inputs of a step (newton Control, control.py:31-110), indexed by global dof; null
// pointers stand for the model's targets (tq) or zero.
struct NaControl {
  const float* jf;   // joint_f: generalized forces; a FREE joint's six are a COM wrench, world frame
  const float* tq;   // joint_target_q (dof layout), replaces the model's joint_target_pos
  const float* tqd;  // joint_target_qd
  const float* act;  // joint_act: feedforward generalized forces
};

struct ArtWorld {
  int w, b0, nb, nl, nseg, q0, nq, d0, nv, dmax;
};

struct OpMaxI {
  __device__ __forceinline__ int operator()(int a, int b) const { return a > b ? a : b; }
};

// ---------------------------------------------------------------- FP64 rigid transforms
struct rtd {
  double p[3], R[9];
};
__device__ __forceinline__ void dqmat(double x, double y, double z, double w, double* R) {
  R[0] = 1.0 - 2.0 * (y * y + z * z); R[1] = 2.0 * (x * y - z * w);       R[2] = 2.0 * (x * z + y * w);
  R[3] = 2.0 * (x * y + z * w);       R[4] = 1.0 - 2.0 * (x * x + z * z); R[5] = 2.0 * (y * z - x * w);
  R[6] = 2.0 * (x * z - y * w);       R[7] = 2.0 * (y * z + x * w);       R[8] = 1.0 - 2.0 * (x * x + y * y);
}
template <typename T>
__device__ __forceinline__ rtd rtd_from7(const T* x) {
  rtd t;
  dqmat((double)x[3], (double)x[4], (double)x[5], (double)x[6], t.R);
  t.p[0] = (double)x[0]; t.p[1] = (double)x[1]; t.p[2] = (double)x[2];
  return t;
}
__device__ __forceinline__ rtd rtd_mul(const rtd& a, const rtd& b) {
  rtd t;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
      t.R[3 * i + j] = a.R[3 * i] * b.R[j] + a.R[3 * i + 1] * b.R[3 + j] + a.R[3 * i + 2] * b.R[6 + j];
    t.p[i] = a.p[i] + a.R[3 * i] * b.p[0] + a.R[3 * i + 1] * b.p[1] + a.R[3 * i + 2] * b.p[2];
  }
  return t;
}
__device__ __forceinline__ rtd rtd_inv(const rtd& a) {
  rtd t;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t.R[3 * i + j] = a.R[3 * j + i];
#pragma unroll
  for (int i = 0; i < 3; ++i) t.p[i] = -(t.R[3 * i] * a.p[0] + t.R[3 * i + 1] * a.p[1] + t.R[3 * i + 2] * a.p[2]);
  return t;
}
// rotation by angle th about the (normalised) axis a
__device__ __forceinline__ rtd rtd_axis_angle(const float* a, double th) {
  const double ax = a[0], ay = a[1], az = a[2];
  const double n = sqrt(ax * ax + ay * ay + az * az);
  const double s = sin(0.5 * th), c = cos(0.5 * th);
  rtd t;
  dqmat(ax / n * s, ay / n * s, az / n * s, c, t.R);
  t.p[0] = t.p[1] = t.p[2] = 0.0;
  return t;
}
// unit quaternion (x, y, z, w) of a rotation matrix (Shepperd's largest-component branch)
__device__ __forceinline__ void dquat_of(const double* R, double* q) {
  const double tr = R[0] + R[4] + R[8];
  if (tr > 0.0) {
    const double s = sqrt(tr + 1.0) * 2.0;
    q[3] = 0.25 * s; q[0] = (R[7] - R[5]) / s; q[1] = (R[2] - R[6]) / s; q[2] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    const double s = sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
    q[3] = (R[7] - R[5]) / s; q[0] = 0.25 * s; q[1] = (R[1] + R[3]) / s; q[2] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    const double s = sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
    q[3] = (R[2] - R[6]) / s; q[0] = (R[1] + R[3]) / s; q[1] = 0.25 * s; q[2] = (R[5] + R[7]) / s;
  } else {
    const double s = sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
    q[3] = (R[3] - R[1]) / s; q[0] = (R[2] + R[6]) / s; q[1] = (R[5] + R[7]) / s; q[2] = 0.25 * s;
  }
}
__device__ __forceinline__ void dcross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
// spatial inertia of mass m, COM c, rotational inertia I about c (sym xx yy zz xy xz yz) applied
// to the motion x = (w; v) at the origin: h = m (v + w x c), result (I w + c x h; h)
__device__ __forceinline__ void sp_inertia(double m, const double* c, const double* I, const double* x, double* o) {
  double wc[3];
  dcross3(x, c, wc);
  const double h[3] = {m * (x[3] + wc[0]), m * (x[4] + wc[1]), m * (x[5] + wc[2])};
  double ch[3];
  dcross3(c, h, ch);
  o[0] = I[0] * x[0] + I[3] * x[1] + I[4] * x[2] + ch[0];
  o[1] = I[3] * x[0] + I[1] * x[1] + I[5] * x[2] + ch[1];
  o[2] = I[4] * x[0] + I[5] * x[1] + I[2] * x[2] + ch[2];
  o[3] = h[0]; o[4] = h[1]; o[5] = h[2];
}

// is dof k's joint on the path from the root to body b?
__device__ __forceinline__ bool art_on_path(const ArtLDS& A, int k, int b) {
  return (A.anc[b] >> A.dof_body[k]) & 1u;
}

// ---------------------------------------------------------------- world set-up
__device__ __forceinline__ ArtWorld art_world(const NaModel& M, int w) {
  ArtWorld a;
  a.w = w;
  a.b0 = M.world_body_start[w];
  a.nb = M.world_body_end[w] - a.b0;
  a.nl = M.world_shape_end[w] - M.world_shape_start[w];
  a.nseg = a.nl + M.G;
  a.q0 = a.nb > 0 ? M.joint_q_start[a.b0] : 0;
  a.d0 = a.nb > 0 ? M.joint_qd_start[a.b0] : 0;
  a.nq = 0;  // set by art_load from the joint types: a world's coordinates and dofs are contiguous
  a.nv = 0;  // from q0 / d0, whatever follows them (a pool keeps worlds in arena ranges)
  a.dmax = 0;
  return a;
}

// Joint data and control inputs of the world into LDS; sets aw.dmax (depth of the deepest body,
// root = 0).  body_f [B,6] (may be null): applied COM wrenches, world frame (State.body_f).
__device__ void art_load(const NaModel& M, ArtWorld& aw, ArtLDS& A, const NaControl& U, const float* body_f) {
  const int l = lane_id();
  if (l < aw.nb) {
    const int j = aw.b0 + l;
    A.jtype[l] = M.joint_type[j];
    const int p = M.joint_parent[j];
    A.jpar[l] = p < 0 ? -1 : p - aw.b0;
    A.jq[l] = M.joint_q_start[j] - aw.q0;
    A.jd[l] = M.joint_qd_start[j] - aw.d0;
#pragma unroll
    for (int k = 0; k < 7; ++k) { A.Xp[l][k] = M.joint_X_p[7 * j + k]; A.Xc[l][k] = M.joint_X_c[7 * j + k]; }
#pragma unroll
    for (int k = 0; k < 3; ++k) { A.axis[l][k] = M.joint_axis[3 * j + k]; A.com_b[l][k] = M.body_com[3 * j + k]; }
  }
  if (l == NA_ART_DUMMY) A.anc[l] = 0u;
  if (l < 6) A.V[NA_ART_DUMMY][l] = 0.0;
  if (l < NA_ART_ND) A.x[l] = 0.0;
  wave_sync();
  int depth = 0;
  if (l < aw.nb) {
    uint32_t m = 0u;
    int x = l;
    for (int g = 0; g < NA_ART_NB && x >= 0; ++g) {  // parents precede children: at most nb hops
      m |= 1u << x;
      ++depth;
      x = A.jpar[x];
    }
    A.anc[l] = m;
    depth -= 1;
    A.depth[l] = depth;
    const bool free_j = A.jtype[l] == NA_JT_FREE;
    const int nd = free_j ? 6 : 1;
    for (int k = 0; k < nd; ++k) {
      A.dof_body[A.jd[l] + k] = l;
      A.dof_q[A.jd[l] + k] = free_j ? -1 : A.jq[l];
    }
  }
  aw.dmax = wave_reduce_i(depth, OpMaxI());
  int jq = 0, jd = 0;
  if (l < aw.nb) {
    const bool free_j = A.jtype[l] == NA_JT_FREE;
    jq = free_j ? 7 : 1;
    jd = free_j ? 6 : 1;
  }
  aw.nq = wave_sum_i(jq);
  aw.nv = wave_sum_i(jd);
  wave_sync();  // dof_body of every lane is written
  if (l < aw.nv) {
    const int d = aw.d0 + l;
    A.kp[l] = M.joint_target_ke[d];
    A.kd[l] = M.joint_target_kd[d];
    A.arm[l] = M.joint_armature[d];
    A.qt[l] = U.tq ? U.tq[d] : M.joint_target_pos[d];
    A.vt[l] = U.tqd ? U.tqd[d] : 0.f;
    // a FREE joint's joint_f is a COM wrench (below); a revolute dof's is a torque
    const bool rev = A.jtype[A.dof_body[l]] != NA_JT_FREE;
    A.tau[l] = (U.act ? U.act[d] : 0.f) + (U.jf && rev ? U.jf[d] : 0.f);
  }
  if (l < aw.nb) {
    const int j = aw.b0 + l;
    const bool free_j = A.jtype[l] == NA_JT_FREE;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float fe = body_f ? body_f[6 * j + k] : 0.f;
      if (free_j && U.jf) fe += U.jf[aw.d0 + A.jd[l] + k];
      A.fext[l][k] = fe;
    }
  }
  wave_sync();
}

// Forward kinematics at the coordinates held by lanes k < nq (`q`): poses, COMs, world
// inertias, motion subspace columns and the f32 collision poses, one depth level at a time.
__device__ void art_kinematics(const NaModel& M, const ArtWorld& aw, lds_ptr_t<ArtLDS> Ap, double q) {
  ArtLDS& A = lds_deref<ArtLDS>(Ap);
  const int l = lane_id();
  if (l < aw.nq) A.qe[l] = q;
  wave_sync();
  const bool own = l < aw.nb;
  for (int d = 0; d <= aw.dmax; ++d) {
    if (own && A.depth[l] == d) {
      const int qo = A.jq[l], dj = A.jd[l];
      rtd X;
      if (A.jtype[l] == NA_JT_FREE) {
        X = rtd_from7<double>(&A.qe[qo]);
        double c[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          c[i] = X.p[i] + X.R[3 * i] * A.com_b[l][0] + X.R[3 * i + 1] * A.com_b[l][1] + X.R[3 * i + 2] * A.com_b[l][2];
        // dofs (v_com, w): columns (0; e_k) and (e_k; c x e_k), since v_origin = v_com + c x w
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double e[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
          double ce[3];
          dcross3(c, e, ce);
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            A.S[dj + k][i] = 0.0;
            A.S[dj + k][3 + i] = e[i];
            A.S[dj + 3 + k][i] = e[i];
            A.S[dj + 3 + k][3 + i] = ce[i];
          }
        }
      } else {
        rtd Xj = rtd_from7<float>(A.Xp[l]);
        const int p = A.jpar[l];
        if (p >= 0) {
          rtd P;
#pragma unroll
          for (int i = 0; i < 9; ++i) P.R[i] = A.Rb[p][i];
#pragma unroll
          for (int i = 0; i < 3; ++i) P.p[i] = A.pb[p][i];
          Xj = rtd_mul(P, Xj);
        }
        X = rtd_mul(rtd_mul(Xj, rtd_axis_angle(A.axis[l], A.qe[qo])), rtd_inv(rtd_from7<float>(A.Xc[l])));
        const double ax = A.axis[l][0], ay = A.axis[l][1], az = A.axis[l][2];
        const double an = sqrt(ax * ax + ay * ay + az * az);
        double a[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) a[i] = (Xj.R[3 * i] * ax + Xj.R[3 * i + 1] * ay + Xj.R[3 * i + 2] * az) / an;
        double oa[3];
        dcross3(Xj.p, a, oa);
#pragma unroll
        for (int i = 0; i < 3; ++i) { A.S[dj][i] = a[i]; A.S[dj][3 + i] = oa[i]; }
      }
#pragma unroll
      for (int i = 0; i < 9; ++i) A.Rb[l][i] = X.R[i];
#pragma unroll
      for (int i = 0; i < 3; ++i) A.pb[l][i] = X.p[i];
    }
    wave_sync();
  }
  if (own) {
    const int gb = aw.b0 + l;
    double R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = A.Rb[l][i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      A.cb[l][i] = A.pb[l][i] + R[3 * i] * A.com_b[l][0] + R[3 * i + 1] * A.com_b[l][1] + R[3 * i + 2] * A.com_b[l][2];
    const float* Ibody = M.body_inertia + 9 * gb;
    double Iw[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int m = 0; m < 3; ++m) s += R[3 * i + k] * (double)Ibody[3 * k + m] * R[3 * j + m];
        Iw[3 * i + j] = s;
      }
    A.Ib[l][0] = Iw[0]; A.Ib[l][1] = Iw[4]; A.Ib[l][2] = Iw[8];
    A.Ib[l][3] = Iw[1]; A.Ib[l][4] = Iw[2]; A.Ib[l][5] = Iw[5];
    A.mb[l] = (double)M.body_mass[gb];
    double qq[4];
    dquat_of(R, qq);
#pragma unroll
    for (int i = 0; i < 3; ++i) A.bqf[i][l] = (float)A.pb[l][i];
#pragma unroll
    for (int i = 0; i < 4; ++i) A.bqf[3 + i][l] = (float)qq[i];
  }
  wave_sync();
}

// Body velocities V_b = sum over the dofs k on body b's path of S_k x_k, from A.x.
__device__ __forceinline__ void art_body_velocity(const ArtWorld& aw, ArtLDS& A) {
  const int l = lane_id();
  if (l < aw.nb) {
    double V[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t an = A.anc[l];
#pragma unroll 4
    for (int k = 0; k < aw.nv; ++k) {
      if (!((an >> A.dof_body[k]) & 1u)) continue;
      const double xk = A.x[k];
#pragma unroll
      for (int i = 0; i < 6; ++i) V[i] += A.S[k][i] * xk;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) A.V[l][i] = V[i];
  }
  wave_sync();
}

// (A.A x)_l for lane l < nv with x from A.x (FP64).
__device__ __forceinline__ double art_matvec(const ArtWorld& aw, const ArtLDS& A) {
  const int l = lane_id();
  double y = 0.0;
  if (l < aw.nv) {
#pragma unroll 8
    for (int j = 0; j < aw.nv; ++j) y += A.A[l][j] * A.x[j];
  }
  return y;
}

// Bias c (returned to lane k < nv) and mass matrix (into A.A; its diagonal returned in `mdiag`)
// at the kinematics in LDS and this lane's dof velocity v.
__device__ double art_dynamics(const NaModel& M, const ArtWorld& aw, ArtLDS& A, double v, double& mdiag) {
  const int l = lane_id();
  const int nb = aw.nb, nv = aw.nv;
  if (l < nv) A.x[l] = v;
  wave_sync();
  art_body_velocity(aw, A);
  // velocity products per dof: revolute (V_b x S_k) v_k; the FREE columns c x e_k move with the
  // COM, which gives (0; v_com x w) on the first FREE dof
  if (l < nv) {
    const int b = A.dof_body[l];
    double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (A.jtype[b] == NA_JT_FREE) {
      const int dj = A.jd[b];
      if (l == dj) {
        const double vc[3] = {A.x[dj], A.x[dj + 1], A.x[dj + 2]};
        const double w[3] = {A.x[dj + 3], A.x[dj + 4], A.x[dj + 5]};
        dcross3(vc, w, o + 3);
      }
    } else {
      double Vb[6], s[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) { Vb[i] = A.V[b][i]; s[i] = A.S[l][i]; }
      double t0[3], t1[3], t2[3];
      dcross3(Vb, s, t0);
      dcross3(Vb, s + 3, t1);
      dcross3(Vb + 3, s, t2);
#pragma unroll
      for (int i = 0; i < 3; ++i) { o[i] = t0[i] * v; o[3 + i] = (t1[i] + t2[i]) * v; }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) A.T[l][i] = o[i];
  }
  wave_sync();
  // body force f_b = I_b a_b + V_b x* (I_b V_b) - (c_b x m g; m g), a_b the sum of the products on
  // the body's path; each lane replaces its own body velocity by the force
  if (l < nb) {
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, Vb[6], c[3], I[6];
    const uint32_t an = A.anc[l];
    for (int k = 0; k < nv; ++k) {
      if (!((an >> A.dof_body[k]) & 1u)) continue;
#pragma unroll
      for (int i = 0; i < 6; ++i) a[i] += A.T[k][i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) { Vb[i] = A.V[l][i]; I[i] = A.Ib[l][i]; }
#pragma unroll
    for (int i = 0; i < 3; ++i) c[i] = A.cb[l][i];
    const double m = A.mb[l];
    double f[6], IV[6];
    sp_inertia(m, c, I, a, f);
    sp_inertia(m, c, I, Vb, IV);
    // V x* F = (w x n + v x f; w x f) for V = (w; v), F = (n; f)
    double t0[3], t1[3], t2[3];
    dcross3(Vb, IV, t0);
    dcross3(Vb + 3, IV + 3, t1);
    dcross3(Vb, IV + 3, t2);
    // gravity and the applied COM wrench (fe; te) act at the COM: (c x f + t; f) at the origin
    const double fe[3] = {m * (double)M.gx + (double)A.fext[l][0], m * (double)M.gy + (double)A.fext[l][1],
                          m * (double)M.gz + (double)A.fext[l][2]};
    double cfe[3];
    dcross3(c, fe, cfe);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      f[i] += t0[i] + t1[i] - cfe[i] - (double)A.fext[l][3 + i];
      f[3 + i] += t2[i] - fe[i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) A.V[l][i] = f[i];
  }
  wave_sync();
  double cbias = 0.0;
  if (l < nv) {
    for (int b = 0; b < nb; ++b) {
      if (!art_on_path(A, l, b)) continue;
#pragma unroll
      for (int i = 0; i < 6; ++i) cbias += A.S[l][i] * A.V[b][i];
    }
  }
  // composite inertia of each body's subtree: mass, COM, rotational inertia about that COM
  double cm = 0.0, cc[3] = {0.0, 0.0, 0.0}, cI[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (l < nb) {
    for (int d = 0; d < nb; ++d) {
      if (!((A.anc[d] >> l) & 1u)) continue;
      const double m = A.mb[d];
      cm += m;
#pragma unroll
      for (int i = 0; i < 3; ++i) cc[i] += m * A.cb[d][i];
    }
    const double icm = cm > 0.0 ? 1.0 / cm : 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) cc[i] *= icm;
    for (int d = 0; d < nb; ++d) {
      if (!((A.anc[d] >> l) & 1u)) continue;
      const double m = A.mb[d];
      const double r[3] = {A.cb[d][0] - cc[0], A.cb[d][1] - cc[1], A.cb[d][2] - cc[2]};
      const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      cI[0] += A.Ib[d][0] + m * (rr - r[0] * r[0]);
      cI[1] += A.Ib[d][1] + m * (rr - r[1] * r[1]);
      cI[2] += A.Ib[d][2] + m * (rr - r[2] * r[2]);
      cI[3] += A.Ib[d][3] - m * r[0] * r[1];
      cI[4] += A.Ib[d][4] - m * r[0] * r[2];
      cI[5] += A.Ib[d][5] - m * r[1] * r[2];
    }
  }
  wave_sync();  // every lane has read the single-body inertias
  if (l < nb) {
    A.mb[l] = cm;
#pragma unroll
    for (int i = 0; i < 3; ++i) A.cb[l][i] = cc[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) A.Ib[l][i] = cI[i];
  }
  wave_sync();
  // F_j = Ic_{body(j)} S_j; column j of M: M_ij = S_i . F_j for i <= j with body(i) on the path of
  // body(j), M_ij = S_j . F_i for i > j with body(j) on the path of body(i), 0 otherwise (exactly
  // symmetric: both triangles take the same products)
  if (l < nv) {
    const int b = A.dof_body[l];
    double s[6], F[6], c[3], I[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) { s[i] = A.S[l][i]; I[i] = A.Ib[b][i]; }
#pragma unroll
    for (int i = 0; i < 3; ++i) c[i] = A.cb[b][i];
    sp_inertia(A.mb[b], c, I, s, F);
#pragma unroll
    for (int i = 0; i < 6; ++i) A.T[l][i] = F[i];
  }
  wave_sync();
  if (l < nv) {
    const int bj = A.dof_body[l];
    for (int i = 0; i < nv; ++i) {
      double val = 0.0;
      if (i <= l) {
        if (art_on_path(A, i, bj)) {
#pragma unroll
          for (int k = 0; k < 6; ++k) val += A.S[i][k] * A.T[l][k];
        }
      } else if (art_on_path(A, l, A.dof_body[i])) {
#pragma unroll
        for (int k = 0; k < 6; ++k) val += A.S[l][k] * A.T[i][k];
      }
      A.A[i][l] = val;
    }
  }
  wave_sync();
  mdiag = l < nv ? A.A[l][l] : 0.0;
  return cbias;
}

// ---------------------------------------------------------------- dense factorisations
// Column broadcasts of the register factorisation (lane k's register j is L[k][j]).
__device__ __forceinline__ float bcast(float v, int k) { return lane_f(v, k); }
__device__ __forceinline__ double bcast(double v, int k) { return lane_d(v, k); }

// Factor the leading n x n block of C (lower triangle) in registers -- lane i owns row i, column j
// of the factor is broadcast with readlane, rows and columns n..NM-1 are padded with the identity
// so that the NM-step sweeps run without guards -- store the factor back to C, and solve
// (L L^T) x = b in FP64: the forward sweep uses each lane's own row of L, the backward sweep its
// column, loaded from LDS in one batch.  Lane i holds b_i and gets x_i; `ok` is false unless every
// pivot is positive.
template <int NM, typename T>
__device__ double art_factor_solve_reg(T (*Cg)[NA_ART_LD], int n, double b, bool& ok) {
  // the matrix lives in LDS: address it as such so that every access is a ds_* instruction
  typedef __attribute__((address_space(3))) T lds_row[NA_ART_LD];
  lds_row* C = (lds_row*)Cg;
  const int l = lane_id();
  n = __builtin_amdgcn_readfirstlane(n);
  // rows and columns n..NM-1 are the identity, so the sweeps below run without guards
  T r[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) r[k] = (l < n && k <= l) ? C[l][k] : (l == k ? (T)1 : (T)0);
  T dg[NM];
  T dmin = (T)1;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    const T djj = bcast(r[j], j);
    dmin = djj < dmin ? djj : dmin;
    const T d = sqrt(djj > (T)0 ? djj : (T)1);
    dg[j] = (T)1 / d;
    r[j] = l == j ? d : r[j] * dg[j];
#pragma unroll
    for (int k = j + 1; k < NM; ++k) r[k] -= r[j] * bcast(r[j], k);
  }
  ok = dmin > (T)0;
  wave_sync();
  if (l < n) {
#pragma unroll
    for (int k = 0; k < NM; ++k)
      if (k <= l) C[l][k] = r[k];
  }
  // forward: L y = b (lane l subtracts L[l][i] y_i with its own row); dg holds 1 / L[i][i]
  double x = l < n ? b : 0.0;
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    const double xi = lane_d(x, i) * (double)dg[i];
    const double upd = x - (double)r[i] * xi;
    x = l == i ? xi : (l > i ? upd : x);
  }
  wave_sync();
  // backward: L^T x = y (lane l subtracts L[i][l] x_i with its column, read from LDS)
  T ct[NM];
#pragma unroll
  for (int i = 0; i < NM; ++i) ct[i] = (i < n && i > l) ? C[i][l] : (T)0;
#pragma unroll
  for (int i = NM - 1; i >= 0; --i) {
    const double xi = lane_d(x, i) * (double)dg[i];
    const double upd = x - (double)ct[i] * xi;
    x = l == i ? xi : (l < i ? upd : x);
  }
  return l < n ? x : 0.0;
}
template <typename T>
__device__ double art_factor_solve(T (*C)[NA_ART_LD], int n, double b, bool& ok) {
  if (n <= 8) return art_factor_solve_reg<8, T>(C, n, b, ok);
  if (n <= 24) return art_factor_solve_reg<24, T>(C, n, b, ok);
  return art_factor_solve_reg<32, T>(C, n, b, ok);
}

// ---------------------------------------------------------------- contacts
// contact frame rows of contact c (t1, t2 = n x t1, n) in FP64
__device__ __forceinline__ void art_frame(const ArtLDS& A, int c, double* F) {
  const double n[3] = {A.con.n[0][c], A.con.n[1][c], A.con.n[2][c]};
  const double t1[3] = {A.t1[0][c], A.t1[1][c], A.t1[2][c]};
  double t2[3];
  dcross3(n, t1, t2);
#pragma unroll
  for (int i = 0; i < 3; ++i) { F[i] = t1[i]; F[3 + i] = t2[i]; F[6 + i] = n[i]; }
}
// velocity of the contact point on body B relative to body A in the contact frame, from A.V
__device__ __forceinline__ void art_point_vel(const ArtLDS& A, int c, const double* F, double* vc) {
  const double p[3] = {A.con.p[0][c], A.con.p[1][c], A.con.p[2][c]};
  const int ba = A.con.ba[c] < 0 ? NA_ART_DUMMY : A.con.ba[c];
  const int bb = A.con.bb[c] < 0 ? NA_ART_DUMMY : A.con.bb[c];
  const double wa[3] = {A.V[ba][0], A.V[ba][1], A.V[ba][2]}, wb[3] = {A.V[bb][0], A.V[bb][1], A.V[bb][2]};
  double xa[3], xb[3];
  dcross3(wa, p, xa);
  dcross3(wb, p, xb);
  double rel[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) rel[i] = (A.V[bb][3 + i] + xb[i]) - (A.V[ba][3 + i] + xa[i]);
#pragma unroll
  for (int r = 0; r < 3; ++r) vc[r] = F[3 * r] * rel[0] + F[3 * r + 1] * rel[1] + F[3 * r + 2] * rel[2];
}
// sign of dof k in the rows of contact c: +1 on body B's path only, -1 on body A's only, else 0
__device__ __forceinline__ double art_sign(const ArtLDS& A, int k, int c) {
  const int ba = A.con.ba[c], bb = A.con.bb[c];
  const double sb = (bb >= 0 && art_on_path(A, k, bb)) ? 1.0 : 0.0;
  const double sa = (ba >= 0 && art_on_path(A, k, ba)) ? 1.0 : 0.0;
  return sb - sa;
}

// |L^-1 J_r^T|^2 for every contact row r of the nc contacts (one row 3c + r per lane, in passes of
// 64), by forward substitution in registers against the factor of A in A.fa.LA; into A.fa.wrow.
template <int NM>
__device__ void art_delassus_rows(const ArtWorld& aw, ArtLDS& A, int nc) {
  const int l = lane_id();
  const int nv = aw.nv;
  for (int base = 0; base < 3 * nc; base += NA_WAVE) {
    const int row = base + l;
    const int c = row / 3, r = row - 3 * c;
    const bool valid = row < 3 * nc;
    double F[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double p[3] = {0.0, 0.0, 0.0};
    if (valid) {
      art_frame(A, c, F);
      p[0] = A.con.p[0][c]; p[1] = A.con.p[1][c]; p[2] = A.con.p[2][c];
    }
    const double fr[3] = {F[3 * r], F[3 * r + 1], F[3 * r + 2]};
    double y[NM];
    double ysq = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
      y[i] = 0.0;
      if (i < nv) {
        // J_r,i = sign_i F_r . (S_i,lin + S_i,ang x p)
        const double sg = valid ? art_sign(A, i, c) : 0.0;
        const double wa[3] = {A.S[i][0], A.S[i][1], A.S[i][2]};
        double u[3];
        dcross3(wa, p, u);
        double ji = sg * (fr[0] * (A.S[i][3] + u[0]) + fr[1] * (A.S[i][4] + u[1]) + fr[2] * (A.S[i][5] + u[2]));
#pragma unroll
        for (int k = 0; k < i; ++k) ji -= A.fa.LA[i][k] * y[k];
        y[i] = ji / A.fa.LA[i][i];
        ysq += y[i] * y[i];
      }
    }
    if (valid) A.fa.wrow[row] = ysq;
  }
  wave_sync();
}

// Per-contact set-up of a step h (lane c < nc): frame, Delassus diagonal through the factor of
// A in A.fa.LA, regularisation, v_hat, activity.
__device__ void art_contact_setup(const NaConfig& cfg, const ArtWorld& aw, lds_ptr_t<ArtLDS> Ap, int nc, double h) {
  ArtLDS& A = lds_deref<ArtLDS>(Ap);
  const int c = lane_id();
  const int nv = aw.nv;
  if (c < nc) {
    const f3 n = mk(A.con.n[0][c], A.con.n[1][c], A.con.n[2][c]);
    const f3 e = fabsf(n.x) < fabsf(n.y) ? (fabsf(n.x) < fabsf(n.z) ? mk(1.f, 0.f, 0.f) : mk(0.f, 0.f, 1.f))
                                         : (fabsf(n.y) < fabsf(n.z) ? mk(0.f, 1.f, 0.f) : mk(0.f, 0.f, 1.f));
    const f3 t1 = nrm3(cross3(n, e));
    A.t1[0][c] = t1.x; A.t1[1][c] = t1.y; A.t1[2][c] = t1.z;
  }
  wave_sync();
  if (nv <= 8) art_delassus_rows<8>(aw, A, nc);
  else if (nv <= 24) art_delassus_rows<24>(aw, A, nc);
  else art_delassus_rows<32>(aw, A, nc);
  if (c < nc) {
    const double wsum = A.fa.wrow[3 * c] + A.fa.wrow[3 * c + 1] + A.fa.wrow[3 * c + 2];
    const double wd = wsum / 3.0;
    const double tau_d = cfg.tau_d;
    if (wd < NA_ART_W_MIN) {
      A.act[c] = 0;
      A.Rn[c] = 1.0; A.Rt[c] = 1.0; A.iRn[c] = 1.0; A.iRt[c] = 1.0; A.vhat[c] = 0.0;
    } else {
      const double rn_soft = 1.0 / (h * (double)A.con.k[c] * (h + tau_d));
      const double rn_rigid = (double)cfg.beta * (double)cfg.beta / (4.0 * 9.8696044010893586188) * wd;
      const double Rn = rn_rigid > rn_soft ? rn_rigid : rn_soft;
      const double Rt = (double)cfg.sigma * wd;
      A.act[c] = 1;
      A.Rn[c] = Rn; A.Rt[c] = Rt; A.iRn[c] = 1.0 / Rn; A.iRt[c] = 1.0 / Rt;
      A.vhat[c] = -(double)A.con.phi0[c] / (h + tau_d);
    }
  }
  wave_sync();
}

// Contact Hessian J^T G J on MFMA plus A (rounded to f32) into the leading nv x nv block of A.H.
// Lane l fills column l % 32 of the rows of contacts 2t + l / 32 of each 16-contact chunk.
__device__ void art_hessian(const ArtWorld& aw, ArtLDS& A, int nc) {
  const int l = lane_id();
  const int k = l & 31;
  const int nv = aw.nv;
  const bool kv = k < nv;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (kv) {
#pragma unroll
    for (int i = 0; i < 6; ++i) s[i] = A.S[k][i];
  }
  na_v16f acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int c0 = 0; c0 < nc; c0 += NA_ART_CC) {
    wave_sync();  // the previous chunk's operands are consumed
    for (int t = 0; t < NA_ART_CC / 2; ++t) {
      const int cl = 2 * t + (l >> 5);
      const int c = c0 + cl;
      float j[3] = {0.f, 0.f, 0.f};
      float gj[3] = {0.f, 0.f, 0.f};
      if (kv && c < nc && A.act[c]) {
        const double sg = art_sign(A, k, c);
        if (sg != 0.0) {
          double F[9];
          art_frame(A, c, F);
          const double p[3] = {A.con.p[0][c], A.con.p[1][c], A.con.p[2][c]};
          double u[3];
          dcross3(s, p, u);
#pragma unroll
          for (int r = 0; r < 3; ++r)
            j[r] = (float)(sg * (F[3 * r] * (s[3] + u[0]) + F[3 * r + 1] * (s[4] + u[1]) + F[3 * r + 2] * (s[5] + u[2])));
          const float g00 = A.G[0][c], g11 = A.G[1][c], g22 = A.G[2][c];
          const float g01 = A.G[3][c], g02 = A.G[4][c], g12 = A.G[5][c];
          gj[0] = g00 * j[0] + g01 * j[1] + g02 * j[2];
          gj[1] = g01 * j[0] + g11 * j[1] + g12 * j[2];
          gj[2] = g02 * j[0] + g12 * j[1] + g22 * j[2];
        }
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        A.mm.P[3 * cl + r][k] = j[r];
        A.mm.Q[3 * cl + r][k] = gj[r];
      }
    }
    wave_sync();
    // C = P^T Q: A operand P[kk][i] and B operand Q[kk][j] with i, j = lane % 32, kk = k0 + lane / 32
#pragma unroll 4
    for (int k0 = 0; k0 < 3 * NA_ART_CC; k0 += 2) {
      const int kk = k0 + (l >> 5);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(A.mm.P[kk][k], A.mm.Q[kk][k], acc, 0, 0, 0);
    }
  }
  // register r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4 of column l % 32
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    if (i < nv && kv) A.H[i][k] = (float)A.A[i][k] + acc[r];
  }
  wave_sync();
}

// ---------------------------------------------------------------- one step
struct ArtStats {
  int iterations, factorizations, reason;
  bool converged;
};

// One convex step of size h from the kinematics and dynamics in LDS (coordinates A.qe, bias
// `cbias`, mass diagonal `mdiag`, start velocity v0; one value per dof lane) with the nc contacts
// in A.con: Newton from `guess`.  The velocity comes back in `vout` (lane k < nv), the integrated
// coordinates in `qout` (lane k < nq).  Returns true when the SAP optimality criterion holds.
__device__ bool art_step(const NaConfig& cfg, const ArtWorld& aw, lds_ptr_t<ArtLDS> Ap, int nc, double h, double v0,
                         double cbias, double mdiag, double guess, double& vout, double& qout, ArtStats& st,
                         long long* wk) {
  ArtLDS& A = lds_deref<ArtLDS>(Ap);
  const int l = lane_id();
  NA_TIC(t_setup);
  const int nv = aw.nv;
  const bool dof = l < nv;
  st.iterations = 0;
  st.factorizations = 0;
  st.reason = NA_EXIT_MAX_ITER;
  st.converged = false;
  // right-hand side (M + diag(armature)) v0 + h (ke (q_t - q) - c); A = M + diag(armature + h (kd + h ke))
  wave_sync();
  if (dof) {
    A.A[l][l] = mdiag;
    A.x[l] = v0;
  }
  wave_sync();
  double rhs = art_matvec(aw, A);
  double diag = 0.0;
  if (dof) {
    const double arm = A.arm[l], ke = A.kp[l], kd = A.kd[l];
    diag = mdiag + (arm + h * (kd + h * ke));
    rhs += arm * v0;
    const int qi = A.dof_q[l];
    if (qi >= 0) rhs += h * ke * ((double)A.qt[l] - A.qe[qi]);
    rhs += h * (kd * (double)A.vt[l] + (double)A.tau[l]);
    rhs -= h * cbias;
  }
  wave_sync();
  if (dof) {
    A.A[l][l] = diag;
    for (int j = 0; j < nv; ++j) A.fa.LA[l][j] = j == l ? diag : A.A[l][j];
  }
  // FP64 factor of A for v* and the Delassus diagonals
  bool ok;
  const double vstar = art_factor_solve<double>(A.fa.LA, nv, rhs, ok);
  wave_sync();
  art_contact_setup(cfg, aw, lds_ptr(A), nc, h);
  const double Dk = dof ? 1.0 / sqrt(diag) : 0.0;
  if (dof) A.x[l] = vstar;
  wave_sync();
  const double Avs = art_matvec(aw, A);
  NA_TOC(wk[7], t_setup);
  const int max_iterations = cfg.max_iterations;
  const double opt_abs = cfg.opt_abs_tol, opt_rel = cfg.opt_rel_tol, ls_rel = cfg.ls_rel_tol;
  double v = dof ? guess : 0.0;
  if (!ok) st.reason = NA_EXIT_CHOL;
  for (int it = 0; ok && it <= max_iterations; ++it) {
    // gradient g = A (v - v*) - J^T gamma in FP64
    NA_TIC(t_eval);
    wave_sync();
    if (dof) A.x[l] = v;
    wave_sync();
    art_body_velocity(aw, A);
    double ccost = 0.0;
    if (l < nc) {
      double w6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      double vc[3] = {0.0, 0.0, 0.0};
      float Gc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (A.act[l]) {
        double F[9];
        art_frame(A, l, F);
        art_point_vel(A, l, F, vc);
        const double iRt = A.iRt[l], iRn = A.iRn[l];
        const d3 y = d3{-vc[0] * iRt, -vc[1] * iRt, -(vc[2] - A.vhat[l]) * iRn};
        d3 gam;
        ccost = sap_project((double)A.con.mu[l], A.Rt[l], A.Rn[l], iRt, iRn, y, gam, Gc);
        double f[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) f[i] = F[i] * gam.x + F[3 + i] * gam.y + F[6 + i] * gam.z;
        const double p[3] = {A.con.p[0][l], A.con.p[1][l], A.con.p[2][l]};
        dcross3(p, f, w6);
        w6[3] = f[0]; w6[4] = f[1]; w6[5] = f[2];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) A.vc0[i][l] = vc[i];
#pragma unroll
      for (int i = 0; i < 6; ++i) { A.W[i][l] = w6[i]; A.G[i][l] = Gc[i]; }
    }
    wave_sync();
    double jt = 0.0, d = 0.0;
    if (dof) {
      double sl[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) sl[i] = A.S[l][i];
#pragma unroll 2
      for (int c = 0; c < nc; ++c) {
        const double sg = art_sign(A, l, c);
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < 6; ++i) t += sl[i] * A.W[i][c];
        jt += sg * t;
      }
      d = v - vstar;
      A.x[l] = d;
    }
    wave_sync();
    const double Ad = art_matvec(aw, A);
    double mom = 0.0, gn = 0.0, pn = 0.0, jn = 0.0, g = 0.0;
    if (dof) {
      g = Ad - jt;
      mom = 0.5 * d * Ad;
      const double p = Ad + Avs;
      gn = (Dk * g) * (Dk * g);
      pn = (Dk * p) * (Dk * p);
      jn = (Dk * jt) * (Dk * jt);
    }
    gn = sqrt(wave_sum_d(gn));
    pn = sqrt(wave_sum_d(pn));
    jn = sqrt(wave_sum_d(jn));
    const double mom0 = wave_sum_d(mom);
    const double l0 = mom0 + wave_sum_d(ccost);
    NA_TOC(wk[8], t_eval);
    if (!(gn == gn)) {
      st.reason = NA_EXIT_NAN;
      ok = false;
      break;
    }
    if (gn <= opt_abs + opt_rel * fmax(pn, jn)) {
      st.converged = true;
      st.reason = NA_EXIT_CONVERGED;
      break;
    }
    if (it == max_iterations) break;

    // Newton direction from H = A + J^T G J (f32 factor, FP64 substitution)
    NA_TIC(t_hess);
    art_hessian(aw, A, nc);
    NA_TOC(wk[9], t_hess);
    NA_TIC(t_fac);
    bool spd;
    const double dv = art_factor_solve<float>(A.H, nv, -g, spd);
    NA_TOC(wk[2], t_fac);  // diagnostic builds: factor + solve cycles in the (unused) CG slot
    if (!spd) {
      st.reason = NA_EXIT_CHOL;
      ok = false;
      break;
    }
    st.factorizations++;
    NA_TIC(t_ls);

    // exact line search along dv: l(a) = mom0 + a b0 + a^2 aa / 2 + contacts(a)
    wave_sync();
    if (dof) A.x[l] = dv;
    wave_sync();
    art_body_velocity(aw, A);
    double lc0 = 0.0, d1c0 = 0.0, d2c0 = 0.0, lc1 = 0.0, d1c1 = 0.0;
    if (l < nc) {
      double u[3] = {0.0, 0.0, 0.0};
      if (A.act[l]) {
        double F[9];
        art_frame(A, l, F);
        art_point_vel(A, l, F, u);
        const double iRt = A.iRt[l], iRn = A.iRn[l], Rt = A.Rt[l], Rn = A.Rn[l], mu = A.con.mu[l];
        const double vh = A.vhat[l];
        const d3 uu = d3{u[0], u[1], u[2]};
        const d3 y0 = d3{-A.vc0[0][l] * iRt, -A.vc0[1][l] * iRt, -(A.vc0[2][l] - vh) * iRn};
        const d3 y1 = d3{-(A.vc0[0][l] + u[0]) * iRt, -(A.vc0[1][l] + u[1]) * iRt,
                         -(A.vc0[2][l] + u[2] - vh) * iRn};
        double d2c1 = 0.0;
        lc0 = sap_line(mu, Rt, Rn, iRt, iRn, y0, uu, d1c0, d2c0);
        lc1 = sap_line(mu, Rt, Rn, iRt, iRn, y1, uu, d1c1, d2c1);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) A.dvc[i][l] = u[i];
    }
    const double Adv = art_matvec(aw, A);
    double aa = 0.0, b0 = 0.0;
    if (dof) {
      aa = dv * Adv;
      b0 = dv * Ad;
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
    double alpha;
    // this lane's contact in registers for the search
    const bool lact = l < nc && A.act[l];
    const int lc = lact ? l : 0;
    const d3 lu = d3{A.dvc[0][lc], A.dvc[1][lc], A.dvc[2][lc]};
    const d3 lv = d3{A.vc0[0][lc], A.vc0[1][lc], A.vc0[2][lc] - A.vhat[lc]};
    const double liRt = A.iRt[lc], liRn = A.iRn[lc], lRt = A.Rt[lc], lRn = A.Rn[lc], lmu = A.con.mu[lc];
    const SapLineK lk = sap_line_k(lmu, lRt, lRn, liRt, liRn);
    alpha = -1.0;
    if (cfg.line_search == 1) {
      alpha = armijo_decay(dl0, [&](double a) {
        double cc = 0.0, d1 = 0.0, d2 = 0.0;
        if (lact) {
          const d3 y = d3{-(lv.x + a * lu.x) * liRt, -(lv.y + a * lu.y) * liRt, -(lv.z + a * lu.z) * liRn};
          cc = sap_line(lk, y, lu, d1, d2) - lc0;
        }
        return a * b0 + 0.5 * a * a * aa + wave_sum_d(cc);
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
        double dlc = 0.0, d2c = 0.0;
        if (lact) {
          const d3 y = d3{-(lv.x + alpha * lu.x) * liRt, -(lv.y + alpha * lu.y) * liRt, -(lv.z + alpha * lu.z) * liRn};
          sap_line(lk, y, lu, dlc, d2c);
        }
        const double dl = b0 + alpha * aa + wave_sum_d(dlc);
        const double ddl = aa + wave_sum_d(d2c);
        if (fabs(dl) <= ls_rel * fabs(dl0)) break;
        if (dl < 0.0) lo = alpha;
        else hi = alpha;
        double an = ddl > 0.0 ? alpha - dl / ddl : -1.0;
        if (!(an > lo && an < hi)) an = 0.5 * (lo + hi);
        if (hi - lo <= 1e-14 * hi) { alpha = an; break; }
        alpha = an;
      }
    }
    if (dof) v += alpha * dv;
    st.iterations = it + 1;
    NA_TOC(wk[11], t_ls);
  }
  vout = dof ? v : 0.0;
  // integrate (lane j < nb owns joint j): revolute q += h v; FREE: the COM moves with v_com, the
  // quaternion with 1/2 h (w, 0) q (normalised), origin = COM - R com
  wave_sync();
  if (dof) A.x[l] = vout;
  wave_sync();
  if (l < aw.nb) {
    const int qj = A.jq[l], dj = A.jd[l];
    if (A.jtype[l] == NA_JT_FREE) {
      const double qx = A.qe[qj + 3], qy = A.qe[qj + 4], qz = A.qe[qj + 5], qw = A.qe[qj + 6];
      double R0[9];
      dqmat(qx, qy, qz, qw, R0);
      const double cx = A.com_b[l][0], cy = A.com_b[l][1], cz = A.com_b[l][2];
      double xc[3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
        xc[i] = A.qe[qj + i] + R0[3 * i] * cx + R0[3 * i + 1] * cy + R0[3 * i + 2] * cz + h * A.x[dj + i];
      const double wx = A.x[dj + 3], wy = A.x[dj + 4], wz = A.x[dj + 5];
      const double dx = wy * qz - wz * qy + wx * qw, dy = wz * qx - wx * qz + wy * qw;
      const double dz = wx * qy - wy * qx + wz * qw, dw = -(wx * qx + wy * qy + wz * qz);
      double nx = qx + 0.5 * h * dx, ny = qy + 0.5 * h * dy, nz = qz + 0.5 * h * dz, nw = qw + 0.5 * h * dw;
      const double nl = sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
      nx /= nl; ny /= nl; nz /= nl; nw /= nl;
      double R1[9];
      dqmat(nx, ny, nz, nw, R1);
#pragma unroll
      for (int i = 0; i < 3; ++i) A.qo[qj + i] = xc[i] - (R1[3 * i] * cx + R1[3 * i + 1] * cy + R1[3 * i + 2] * cz);
      A.qo[qj + 3] = nx; A.qo[qj + 4] = ny; A.qo[qj + 5] = nz; A.qo[qj + 6] = nw;
    } else {
      A.qo[qj] = A.qe[qj] + h * A.x[dj];
    }
  }
  wave_sync();
  qout = l < aw.nq ? A.qo[l] : 0.0;
  return ok && st.converged;
}

// Contacts of the world at the kinematics in LDS; returns the kept count (<= the capacity).
__device__ int art_collide(const NaModel& M, const NaConfig& cfg, const ArtWorld& aw, lds_ptr_t<ArtLDS> Ap, int& overflow,
                           long long* wk) {
  ArtLDS& A = lds_deref<ArtLDS>(Ap);
  NA_TIC(t0);
  stage_aabb(M, aw.w, aw.nseg, aw.nl, aw.b0, &A.bqf[0][0], NA_ART_NB, A.seg);
  const int np = stage_broad(M, aw.w, aw.nseg, aw.nl, cfg.broad_phase, A.seg);
  const int cap = min(cfg.max_contacts, NA_ART_NC);
  const int nc = stage_narrow(M, aw.b0, np, cap, A.seg, A.con);
  if (np > NA_MAX_PAIRS || nc > cap) overflow = 1;
  wk[3] += min(nc, cap);
  wk[4] += np;
  NA_TOC(wk[6], t0);
  return min(nc, cap);
}

// Step-doubling attempt of an articulated world from (q, v) (lanes k < nq / nv): one step of dt
// and two of dt / 2, the second half-step with contacts at the midpoint.  Returns the FP64 max
// norm of the coordinate difference (1e10 for NaN / inf); `ok` false when a solve failed.  DUMP
// builds (parity entry point only) also write both passes' contact sets to
// dump[2][NA_MAX_CONTACTS][NA_DUMP_FLOATS] and their counts to dump_n[2].
template <bool DUMP = false>
__device__ double art_attempt(const NaModel& M, const NaConfig& cfg, const ArtWorld& aw, lds_ptr_t<ArtLDS> Ap, double q,
                              double v, double dt, double& qf, double& vf, double& qd, double& vd, bool& ok,
                              int& overflow, int& nc_out, long long* wk, float* dump = nullptr,
                              int32_t* dump_n = nullptr) {
  ArtLDS& A = lds_deref<ArtLDS>(Ap);
  const int l = lane_id();
  ArtStats st;
  double mdiag;
  NA_TIC(t_kin);
  art_kinematics(M, aw, lds_ptr(A), q);
  double cb = art_dynamics(M, aw, A, v, mdiag);
  NA_TOC(wk[10], t_kin);
  int nc = art_collide(M, cfg, aw, lds_ptr(A), overflow, wk);
  nc_out = nc;
  if (DUMP) {
    dump_contacts(A.seg, A.con, aw.b0, nc, dump);
    if (l == 0) dump_n[0] = nc;
  }
  ok = art_step(cfg, aw, lds_ptr(A), nc, dt, v, cb, mdiag, v, vf, qf, st, wk);
  wk[1] += st.iterations;
  wk[5] += 1;
  double err = 0.0;
  if (cfg.mode == 1) {
    const double g = 0.5 * (v + vf);
    double qm, vm;
    ok = art_step(cfg, aw, lds_ptr(A), nc, 0.5 * dt, v, cb, mdiag, g, vm, qm, st, wk) && ok;
    wk[1] += st.iterations;
    NA_TIC(t_kin2);
    art_kinematics(M, aw, lds_ptr(A), qm);
    cb = art_dynamics(M, aw, A, vm, mdiag);
    NA_TOC(wk[10], t_kin2);
    nc = art_collide(M, cfg, aw, lds_ptr(A), overflow, wk);
    if (DUMP) {
      dump_contacts(A.seg, A.con, aw.b0, nc, dump + NA_MAX_CONTACTS * NA_DUMP_FLOATS);
      if (l == 0) dump_n[1] = nc;
    }
    ok = art_step(cfg, aw, lds_ptr(A), nc, 0.5 * dt, vm, cb, mdiag, vf, vd, qd, st, wk) && ok;
    wk[1] += st.iterations;
    wk[5] += 2;
    double e = 0.0;
    if (l < aw.nq) e = fabs(qd - qf);
    err = wave_max_d(e);
  } else {
    qd = qf;
    vd = vf;
    double e = 0.0;
    if (l < aw.nq && qf != qf) e = qf;
    err = wave_max_d(e);
  }
  if (isnan(err) || isinf(err)) err = 1.0e10;
  wk[0] += 1;
  return err;
}
