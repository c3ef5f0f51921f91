// Contact generation of one world inside one wavefront:
//   shape world poses + AABBs (one lane per shape)      collide.py:150-275
//   sweep-and-prune / all-pairs broad phase in LDS      broad_phase_sap.py:144-394, broad_phase_nxn.py:129
//   primitive narrow phase (one lane per pair)          narrow_phase.py:154-604, collision_primitive.py
//   gap check + ballot/scan compaction into LDS slots   contact_data.py:86-117, collide.py:62-147
#pragma once
#include "na_math.h"

#define NA_MAX_BODIES 64
#define NA_MAX_SEG 128
#define NA_MAX_PAIRS 1024
#define NA_MAX_CONTACTS 128
#define NA_MAXVAL 1.0e10f

enum { GEO_PLANE = 1, GEO_SPHERE = 3, GEO_CAPSULE = 4, GEO_BOX = 7 };

// Device view of the model (SoA arrays in HBM).
struct NaModel {
  int W, B, S, G, F;
  const int32_t* world_body_start;
  const int32_t* world_shape_start;
  const int32_t* world_body_end;   // one past the world's last body (world_body_start + 1 for a packed model)
  const int32_t* world_shape_end;  // one past the world's last shape
  const int32_t* global_shapes;
  const float* body_com;          // [B,3]
  const float* body_mass;         // [B]
  const float* body_inv_mass;     // [B]
  const float* body_inertia;      // [B,9] body frame
  const float* body_inv_inertia;  // [B,9]
  const int32_t* shape_type;
  const int32_t* shape_body;
  const int32_t* shape_group;
  const int32_t* shape_world;
  const float* shape_xform;  // [S,7]
  const float* shape_scale;  // [S,3]
  const float* shape_margin;
  const float* shape_gap;
  const float* shape_ke;
  const float* shape_mu;
  const float* shape_radius;
  const int32_t* filter_pairs;  // [F,2] sorted canonical
  int P;                         // explicit candidate pairs (0: none given)
  const int32_t* pair_list;      // [P,2] canonical (min, max) global shape indices, grouped by world
  const int32_t* world_pair_start;  // [W+1] range of each world's pairs in pair_list
  float gx, gy, gz;
  // articulated worlds (null for free bodies): joint j moves body j, parents before children
  const int32_t* joint_type;      // [B] 1 revolute, 4 FREE (root only)
  const int32_t* joint_parent;    // [B] global parent body, -1 world
  const int32_t* joint_q_start;   // [B+1]
  const int32_t* joint_qd_start;  // [B+1]
  const float* joint_X_p;         // [B,7] joint frame in the parent frame
  const float* joint_X_c;         // [B,7] joint frame in the child frame
  const float* joint_axis;        // [B,3] revolute axis in the joint frame
  const float* joint_target_pos;  // [D] per dof
  const float* joint_target_ke;
  const float* joint_target_kd;
  const float* joint_armature;
  // MuJoCo geom parameters per shape (null unless the MuJoCo contact law is used): solref [S,2]
  // (time constant, damping ratio; direct stiffness/damping when the first entry is <= 0),
  // solimp [S,5], solmix [S], priority [S], condim [S] (1 or 3); mj_zero_margins: contacts use
  // the surface distance instead of dist - margins, as SolverMuJoCo zeroes margins of box models
  const float* mj_solref;
  const float* mj_solimp;
  const float* mj_solmix;
  const int32_t* mj_priority;
  const int32_t* mj_condim;
  int mj_zero_margins;
};

// Per-wave LDS scratch of the segment (world shapes + global shapes).
struct SegLDS {
  int gidx[NA_MAX_SEG];
  float pose[7][NA_MAX_SEG];
  float lo[3][NA_MAX_SEG];
  float hi[3][NA_MAX_SEG];
  float proj_lo[NA_MAX_SEG];
  float proj_hi[NA_MAX_SEG];
  int order[NA_MAX_SEG];
  int scan[NA_MAX_SEG];
  uint32_t pairs[NA_MAX_PAIRS];  // (seg_a << 8) | seg_b
};

// Per-wave LDS contact set of capacity N, consumed by the convex solve.
template <int N>
struct ContactSet {
  int8_t ba[N], bb[N];    // local body index or -1 (static)
  uint8_t sa[N], sb[N];   // segment positions (A, B after type ordering)
  uint8_t sub[N];
  float n[3][N];          // unit normal A -> B
  float p[3][N];          // midpoint between the surfaces (world)
  float dist[N];          // surface distance (no margins)
  float phi0[N];          // dist - margin_a - margin_b
  float k[N];
  float mu[N];
  float ra[N], rb[N];     // effective radii (writer core points)
};
using ContactLDS = ContactSet<NA_MAX_CONTACTS>;

// ---------------------------------------------------------------- primitives
// up to 8 contacts per pair: box-box keeps its whole clipped face polygon
#define NA_MF 8
struct Manifold {
  float dist[NA_MF];
  f3 pos[NA_MF];
  f3 nrm[NA_MF];
};
__device__ __forceinline__ void mf_clear(Manifold& m) {
#pragma unroll
  for (int i = 0; i < NA_MF; ++i) m.dist[i] = NA_MAXVAL;
}

__device__ __forceinline__ f3 closest_segment_point(f3 a, f3 b, f3 pt) {
  f3 ab = b - a;
  float t = dot3(pt - a, ab) / (dot3(ab, ab) + 1e-6f);
  return a + ab * clampf(t, 0.f, 1.f);
}
__device__ __forceinline__ void plane_sphere(f3 n, f3 pp, f3 sp, float r, float& d, f3& pos) {
  d = dot3(sp - pp, n) - r;
  pos = sp - n * (r + 0.5f * d);
}
__device__ __forceinline__ void sphere_sphere(f3 p1, float r1, f3 p2, float r2, float& d, f3& pos, f3& n) {
  f3 dir = p2 - p1;
  float l = len3(dir);
  n = (l == 0.f) ? mk(1.f, 0.f, 0.f) : dir * (1.f / l);
  d = l - (r1 + r2);
  pos = p1 + n * (r1 + 0.5f * d);
}
__device__ __forceinline__ void sphere_capsule(f3 sp, float sr, f3 cp, f3 ax, float cr, float hl, float& d, f3& pos,
                                               f3& n) {
  f3 seg = ax * hl;
  f3 pt = closest_segment_point(cp - seg, cp + seg, sp);
  sphere_sphere(sp, sr, pt, cr, d, pos, n);
}
__device__ __forceinline__ void capsule_capsule(f3 p1, f3 a1, float r1, float h1, f3 p2, f3 a2, float r2, float h2, Manifold& m) {
  f3 ax1 = a1 * h1, ax2 = a2 * h2, dif = p1 - p2;
  float ma = dot3(ax1, ax1), mb = -dot3(ax1, ax2), mc = dot3(ax2, ax2);
  float u = -dot3(ax1, dif), v = dot3(ax2, dif);
  float det = ma * mc - mb * mb;
  f3 n0;
  if (fabsf(det) >= 1e-15f) {
    float inv = 1.f / det;
    float x1 = (mc * u - mb * v) * inv, x2 = (ma * v - mb * u) * inv;
    if (x1 > 1.f) { x1 = 1.f; x2 = (v - mb) / mc; }
    else if (x1 < -1.f) { x1 = -1.f; x2 = (v + mb) / mc; }
    if (x2 > 1.f) { x2 = 1.f; x1 = clampf((u - mb) / ma, -1.f, 1.f); }
    else if (x2 < -1.f) { x2 = -1.f; x1 = clampf((u + mb) / ma, -1.f, 1.f); }
    sphere_sphere(p1 + ax1 * x1, r1, p2 + ax2 * x2, r2, m.dist[0], m.pos[0], n0);
  } else {
    float x2 = clampf((v - mb) / mc, -1.f, 1.f);
    sphere_sphere(p1 + ax1, r1, p2 + ax2 * x2, r2, m.dist[0], m.pos[0], n0);
    f3 n1;
    x2 = clampf((v + mb) / mc, -1.f, 1.f);
    sphere_sphere(p1 - ax1, r1, p2 + ax2 * x2, r2, m.dist[1], m.pos[1], n1);
  }
  m.nrm[0] = n0;
  m.nrm[1] = n0;
}
// The deepest four corners within `margin`, in the slots of an insertion that replaces the
// shallowest kept corner.  All manifold indices are compile-time constants after unrolling, so
// the manifold stays in registers.
__device__ __forceinline__ void plane_box(f3 n, f3 pp, f3 bp, const m3& R, f3 h, float margin, Manifold& m) {
  float cd0 = dot3(bp - pp, n);
  int nc = 0, worst = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    f3 c = mk((i & 1) ? h.x : -h.x, (i & 2) ? h.y : -h.y, (i & 4) ? h.z : -h.z);
    c = mv(R, c);
    float cd = cd0 + dot3(n, c);
    if (cd > margin) continue;
    f3 cpos = c + bp - n * (0.5f * cd);
    float dw = m.dist[0];
#pragma unroll
    for (int k = 1; k < 4; ++k)
      if (k == worst) dw = m.dist[k];
    if (nc < 4) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k == nc) { m.dist[k] = cd; m.pos[k] = cpos; }
      if (nc == 0 || cd > dw) worst = nc;
      nc++;
    } else if (cd < dw) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k == worst) { m.dist[k] = cd; m.pos[k] = cpos; }
      worst = 0;
      float wd = m.dist[0];
#pragma unroll
      for (int k = 1; k < 4; ++k)
        if (m.dist[k] > wd) { wd = m.dist[k]; worst = k; }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) m.nrm[k] = n;
}
__device__ void sphere_box(f3 sp, float sr, f3 bp, const m3& R, f3 h, float& d, f3& pos, f3& n) {
  f3 c = mtv(R, sp - bp);
  f3 cl = mk(fmaxf(-h.x, fminf(h.x, c.x)), fmaxf(-h.y, fminf(h.y, c.y)), fmaxf(-h.z, fminf(h.z, c.z)));
  f3 dv = cl - c;
  float l = len3(dv);
  f3 cdir = l > 0.f ? dv * (1.f / l) : dv;
  f3 p, nl;
  if (l <= 1e-6f) {
    float closest = 2.f * (h.x + h.y + h.z);
    int k = 0;
    for (int i = 0; i < 6; ++i) {
      float fd = fabsf(((i & 1) ? 1.f : -1.f) * comp3(h, i >> 1) - comp3(c, i >> 1));
      if (closest > fd) { closest = fd; k = i; }
    }
    nl = mk(0.f, 0.f, 0.f);
    setc3(nl, k >> 1, (k & 1) ? -1.f : 1.f);
    p = c + nl * ((sr - closest) * 0.5f);
    d = -closest - sr;
  } else {
    p = (cl + (c + cdir * sr)) * 0.5f;
    nl = cdir;
    d = l - sr;
  }
  n = mv(R, nl);
  pos = bp + mv(R, p);
}

// signed distance of a box-frame point to the box and its outward gradient
__device__ __forceinline__ float box_sdf(f3 p, f3 h, f3& g) {
  f3 q = mk(fabsf(p.x) - h.x, fabsf(p.y) - h.y, fabsf(p.z) - h.z);
  f3 qo = mk(fmaxf(q.x, 0.f), fmaxf(q.y, 0.f), fmaxf(q.z, 0.f));
  float out = len3(qo);
  float in = fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.f);
  if (out > 0.f) {
    g = mk(p.x < 0.f ? -qo.x : qo.x, p.y < 0.f ? -qo.y : qo.y, p.z < 0.f ? -qo.z : qo.z) * (1.f / out);
  } else {
    int k = (q.x >= q.y && q.x >= q.z) ? 0 : (q.y >= q.z ? 1 : 2);
    g = mk(0.f, 0.f, 0.f);
    setc3(g, k, comp3(p, k) < 0.f ? -1.f : 1.f);
  }
  return out + in;
}
#define NA_CAPBOX_ITERS 32
#define NA_CAPBOX_PARALLEL 0.2f
__device__ __forceinline__ void capsule_box(f3 cp, f3 ax, float r, float hl, f3 bp, const m3& R, f3 h, Manifold& m) {
  f3 p0 = mtv(R, (cp - ax * hl) - bp);
  f3 p1 = mtv(R, (cp + ax * hl) - bp);
  f3 d = p1 - p0;
  // minimiser of the convex sdf(p0 + t d) over [0, 1]: bisection on the sign of the directional
  // derivative grad(sdf).d (accurate to rounding, where comparing values is only ~sqrt(eps))
  f3 g;
  float a = 0.f, b = 1.f;
  box_sdf(p0, h, g);
  if (dot3(g, d) >= 0.f) {
    b = 0.f;
  } else {
    box_sdf(p0 + d, h, g);
    if (dot3(g, d) <= 0.f) {
      a = 1.f;
    } else {
      for (int it = 0; it < NA_CAPBOX_ITERS; ++it) {
        float mid = 0.5f * (a + b);
        box_sdf(p0 + d * mid, h, g);
        if (dot3(g, d) < 0.f) a = mid;
        else b = mid;
      }
    }
  }
  float ts = 0.5f * (a + b);
  f3 gs;
  float fs = box_sdf(p0 + d * ts, h, gs);
  float dl = len3(d);
  // Along a face the distance is linear in t over the part of the segment above the face, so
  // its minimum is at an end of that part (not unique when parallel).  Contacts: the ends of
  // the segment clipped to the face slab, plus the golden-section minimum when it lies outside
  // the clipped part (where it is unique).
  bool par = dl > 0.f && fabsf(dot3(d, gs)) < NA_CAPBOX_PARALLEL * dl;
  float tl[3] = {ts, ts, ts};
  int nt = 1;
  if (par) {
    int fk = (fabsf(gs.x) >= fabsf(gs.y) && fabsf(gs.x) >= fabsf(gs.z)) ? 0 : (fabsf(gs.y) >= fabsf(gs.z) ? 1 : 2);
    float lo = 0.f, hi = 1.f;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j == fk) continue;
      float pj = comp3(p0, j), dj = comp3(d, j), hj = comp3(h, j);
      if (fabsf(dj) < 1e-12f) {
        if (fabsf(pj) > hj) { lo = 1.f; hi = 0.f; }
      } else {
        float t1 = (-hj - pj) / dj, t2 = (hj - pj) / dj;
        lo = fmaxf(lo, fminf(t1, t2));
        hi = fminf(hi, fmaxf(t1, t2));
      }
    }
    if (lo <= hi) {
      tl[0] = lo;
      nt = 1;
      if (hi - lo > 1e-6f) { tl[1] = hi; nt = 2; }
      f3 g_;
      float fe = fminf(box_sdf(p0 + d * lo, h, g_), box_sdf(p0 + d * hi, h, g_));
      // the minimum is strictly deeper than the clipped ends (off the face, or inside the box)
      if (ts < lo - 1e-4f || ts > hi + 1e-4f || fs < fe - 1e-6f) {
        if (nt == 1) tl[1] = ts;
        else tl[2] = ts;
        nt++;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k >= nt) continue;
    f3 pk = p0 + d * tl[k];
    f3 gk;
    float fk = box_sdf(pk, h, gk);
    f3 nl = -gk;
    float dist = fk - r;
    f3 mid = pk + nl * (r + 0.5f * dist);
    m.dist[k] = dist;
    m.pos[k] = bp + mv(R, mid);
    m.nrm[k] = mv(R, nl);
  }
}

#define NA_BB_EDGE_BIAS 1.0e-4f
__device__ __forceinline__ int clip_poly(const f3* in, int n, f3 pn, float off, f3* out) {
  int m = 0;
  for (int i = 0; i < n; ++i) {
    f3 a = in[i], b = in[(i + 1) % n];
    float da = dot3(pn, a) - off, db = dot3(pn, b) - off;
    if (da <= 0.f) out[m++] = a;
    if ((da < 0.f && db > 0.f) || (da > 0.f && db < 0.f)) {
      float t = da / (da - db);
      out[m++] = a + (b - a) * t;
    }
  }
  return m;
}
__device__ __forceinline__ void box_box(f3 pa, const m3& Ra, f3 ha, f3 pb, const m3& Rb, f3 hb, float gap, Manifold& m) {
  f3 dv = pb - pa;
  f3 A[3] = {col3(Ra, 0), col3(Ra, 1), col3(Ra, 2)};
  f3 B[3] = {col3(Rb, 0), col3(Rb, 1), col3(Rb, 2)};
  float hA[3] = {ha.x, ha.y, ha.z}, hB[3] = {hb.x, hb.y, hb.z};
  float best = -NA_MAXVAL;
  int best_axis = -1;
  f3 bn = mk(0.f, 0.f, 1.f);
  float scale = fmaxf(fmaxf(ha.x, fmaxf(ha.y, ha.z)), fmaxf(hb.x, fmaxf(hb.y, hb.z)));
  for (int k = 0; k < 15; ++k) {
    f3 L;
    if (k < 3) L = A[k];
    else if (k < 6) L = B[k - 3];
    else {
      L = cross3(A[(k - 6) / 3], B[(k - 6) % 3]);
      float l = len3(L);
      if (l < 1e-6f) continue;
      L = L * (1.f / l);
    }
    float rA = hA[0] * fabsf(dot3(A[0], L)) + hA[1] * fabsf(dot3(A[1], L)) + hA[2] * fabsf(dot3(A[2], L));
    float rB = hB[0] * fabsf(dot3(B[0], L)) + hB[1] * fabsf(dot3(B[1], L)) + hB[2] * fabsf(dot3(B[2], L));
    float dl = dot3(dv, L);
    float s = fabsf(dl) - rA - rB;
    float bias = k < 6 ? 0.f : NA_BB_EDGE_BIAS * scale;
    if (s > best + bias) {
      best = s;
      best_axis = k;
      bn = dl < 0.f ? -L : L;
    }
  }
  if (best > gap) return;
  f3 n = bn;
  if (best_axis < 6) {
    bool ref_a = best_axis < 3;
    const f3* Rr = ref_a ? A : B;
    const f3* Ii = ref_a ? B : A;
    const float* hR = ref_a ? hA : hB;
    const float* hI = ref_a ? hB : hA;
    f3 pR = ref_a ? pa : pb, pI = ref_a ? pb : pa;
    f3 nref = ref_a ? n : -n;
    int ax = ref_a ? best_axis : best_axis - 3;
    f3 fc = pR + nref * hR[ax];
    int ik = 0;
    float mind = NA_MAXVAL, sg = 1.f;
    for (int k = 0; k < 3; ++k) {
      float dd = dot3(Ii[k], nref);
      if (dd < mind) { mind = dd; ik = k; sg = 1.f; }
      if (-dd < mind) { mind = -dd; ik = k; sg = -1.f; }
    }
    f3 ic = pI + Ii[ik] * (sg * hI[ik]);
    int u = (ik + 1) % 3, w = (ik + 2) % 3;
    f3 poly[8], tmp[8];
    poly[0] = ic + (Ii[u] * hI[u] + Ii[w] * hI[w]);
    poly[1] = ic + (Ii[u] * -hI[u] + Ii[w] * hI[w]);
    poly[2] = ic + (Ii[u] * -hI[u] + Ii[w] * -hI[w]);
    poly[3] = ic + (Ii[u] * hI[u] + Ii[w] * -hI[w]);
    int ru = (ax + 1) % 3, rw = (ax + 2) % 3;
    int np = clip_poly(poly, 4, Rr[ru], dot3(Rr[ru], pR) + hR[ru], tmp);
    np = clip_poly(tmp, np, -Rr[ru], -dot3(Rr[ru], pR) + hR[ru], poly);
    np = clip_poly(poly, np, Rr[rw], dot3(Rr[rw], pR) + hR[rw], tmp);
    np = clip_poly(tmp, np, -Rr[rw], -dot3(Rr[rw], pR) + hR[rw], poly);
    // the whole clipped polygon is the manifold (continuous in the poses; a reduction to 4
    // points would switch between near-equal-depth candidates).  Vertex i goes to slot i, so
    // the kept contacts stay in polygon order and the manifold indices are constants.
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i >= np) continue;
      float di = dot3(poly[i] - fc, nref);
      if (di > gap) continue;
      m.dist[i] = di;
      m.pos[i] = poly[i] - nref * (0.5f * di);
      m.nrm[i] = n;
    }
  } else {
    int ia = (best_axis - 6) / 3, ib = (best_axis - 6) % 3;
    f3 ca = pa, cb = pb;
    for (int k = 0; k < 3; ++k) {
      if (k != ia) ca = ca + A[k] * ((dot3(A[k], n) > 0.f ? 1.f : -1.f) * hA[k]);
      if (k != ib) cb = cb + B[k] * ((dot3(B[k], n) > 0.f ? -1.f : 1.f) * hB[k]);
    }
    f3 da = A[ia], db = B[ib], r = ca - cb;
    float bb = dot3(da, db), c = dot3(da, r), f = dot3(db, r);
    float den = 1.f - bb * bb;
    float s = den > 1e-12f ? (bb * f - c) / den : 0.f;
    s = clampf(s, -hA[ia], hA[ia]);
    float t = clampf(bb * s + f, -hB[ib], hB[ib]);
    s = clampf(bb * t - c, -hA[ia], hA[ia]);
    f3 qa = ca + da * s, qb = cb + db * t;
    m.dist[0] = dot3(qb - qa, n);
    m.pos[0] = (qa + qb) * 0.5f;
    m.nrm[0] = n;
  }
}

// ---------------------------------------------------------------- world-level stages
__device__ __forceinline__ int seg_shape(const NaModel& M, int w, int i, int nl) {
  return i < nl ? M.world_shape_start[w] + i : M.global_shapes[i - nl];
}

// Stage 1: world pose + AABB of every segment shape (collide.py:150-275).  Pose of the
// body comes from `bq` (f32 [nb*7] in LDS or HBM, world-local body indexing).
__device__ void stage_aabb(const NaModel& M, int w, int n, int nl, int b0, const float* __restrict__ bq, int bq_stride,
                           SegLDS& S) {
  for (int i = lane_id(); i < n; i += NA_WAVE) {
    int s = seg_shape(M, w, i, nl);
    S.gidx[i] = s;
    const float* sx = M.shape_xform + 7 * s;
    f3 lp = mk(sx[0], sx[1], sx[2]);
    qf lq = qf{sx[3], sx[4], sx[5], sx[6]};
    int b = M.shape_body[s];
    f3 p = lp;
    qf q = lq;
    if (b >= 0) {
      int lb = b - b0;
      f3 bp = mk(bq[lb], bq[lb + bq_stride], bq[lb + 2 * bq_stride]);
      qf bqq = qf{bq[lb + 3 * bq_stride], bq[lb + 4 * bq_stride], bq[lb + 5 * bq_stride], bq[lb + 6 * bq_stride]};
      p = bp + qrot(bqq, lp);
      q = qmul(bqq, lq);
    }
    S.pose[0][i] = p.x; S.pose[1][i] = p.y; S.pose[2][i] = p.z;
    S.pose[3][i] = q.x; S.pose[4][i] = q.y; S.pose[5][i] = q.z; S.pose[6][i] = q.w;
    m3 R = qmat(q.x, q.y, q.z, q.w);
    const float* sc = M.shape_scale + 3 * s;
    float eg = M.shape_margin[s] + M.shape_gap[s];
    int t = M.shape_type[s];
    f3 h;
    if (t == GEO_PLANE && sc[0] == 0.f && sc[1] == 0.f) {
      float r = M.shape_radius[s];
      h = mk(r, r, r);
    } else if (t == GEO_SPHERE) {
      h = mk(sc[0], sc[0], sc[0]);
    } else if (t == GEO_CAPSULE) {
      f3 ax = col3(R, 2);
      h = mk(fabsf(ax.x) * sc[1] + sc[0], fabsf(ax.y) * sc[1] + sc[0], fabsf(ax.z) * sc[1] + sc[0]);
    } else {
      f3 e = (t == GEO_PLANE) ? mk(0.5f * sc[0], 0.5f * sc[1], 0.f) : mk(sc[0], sc[1], sc[2]);
      h = mk(fabsf(R.m[0]) * e.x + fabsf(R.m[1]) * e.y + fabsf(R.m[2]) * e.z,
             fabsf(R.m[3]) * e.x + fabsf(R.m[4]) * e.y + fabsf(R.m[5]) * e.z,
             fabsf(R.m[6]) * e.x + fabsf(R.m[7]) * e.y + fabsf(R.m[8]) * e.z);
    }
    S.lo[0][i] = p.x - h.x - eg; S.lo[1][i] = p.y - h.y - eg; S.lo[2][i] = p.z - h.z - eg;
    S.hi[0][i] = p.x + h.x + eg; S.hi[1][i] = p.y + h.y + eg; S.hi[2][i] = p.z + h.z + eg;
  }
  wave_sync();
}

__device__ __forceinline__ bool group_pair(int a, int b) {
  if (a == 0 || b == 0) return false;
  if (a > 0) return a == b || b < 0;
  return a != b;
}
__device__ __forceinline__ bool pair_excluded(const NaModel& M, int a, int b) {
  int lo = 0, hi = M.F - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int fa = M.filter_pairs[2 * mid], fb = M.filter_pairs[2 * mid + 1];
    if (fa == a && fb == b) return true;
    if (a < fa || (a == fa && b < fb)) hi = mid - 1;
    else lo = mid + 1;
  }
  return false;
}
// world / group / exclusion / AABB test of segment positions i, j (broad_phase_common.py:18-180)
__device__ __forceinline__ bool pair_passes(const NaModel& M, const SegLDS& S, int i, int j) {
  int a = S.gidx[i], b = S.gidx[j];
  int sa = min(a, b), sb = max(a, b);
  int wa = M.shape_world[sa], wb = M.shape_world[sb];
  if (wa == -1 && wb == -1) return false;
  if (wa != -1 && wb != -1 && wa != wb) return false;
  if (!group_pair(M.shape_group[sa], M.shape_group[sb])) return false;
  int ba = M.shape_body[sa], bb = M.shape_body[sb];
  if (ba == bb) return false;  // shapes of one body never collide; static shapes (-1) are one body
  if (M.F > 0 && pair_excluded(M, sa, sb)) return false;
  return S.lo[0][i] <= S.hi[0][j] && S.hi[0][i] >= S.lo[0][j] && S.lo[1][i] <= S.hi[1][j] &&
         S.hi[1][i] >= S.lo[1][j] && S.lo[2][i] <= S.hi[2][j] && S.hi[2][i] >= S.lo[2][j];
}

// Append pair (i, j) from lanes whose predicate holds, preserving lane order.
__device__ __forceinline__ void append_pair(bool pred, int i, int j, SegLDS& S, int& np) {
  int total;
  int r = ballot_rank(pred, &total);
  if (pred && np + r < NA_MAX_PAIRS) S.pairs[np + r] = ((uint32_t)i << 8) | (uint32_t)j;
  np += total;
}

// Stage 2a: all-pairs broad phase (broad_phase_nxn.py:129): lexicographic (i<j) enumeration.
__device__ int stage_broad_nxn(const NaModel& M, int n, SegLDS& S) {
  int total = n * (n - 1) / 2;
  int np = 0;
  for (int base = 0; base < total; base += NA_WAVE) {
    int k = base + lane_id();
    bool pred = false;
    int i = 0, j = 0;
    if (k < total) {
      // invert k = i*n - i*(i+1)/2 + (j - i - 1)
      float fn = (float)n;
      int ii = (int)floorf((2.f * fn - 1.f - sqrtf((2.f * fn - 1.f) * (2.f * fn - 1.f) - 8.f * (float)k)) * 0.5f);
      ii = max(0, min(ii, n - 2));
      while (ii > 0 && ii * n - ii * (ii + 1) / 2 > k) --ii;
      while (ii + 1 <= n - 2 && (ii + 1) * n - (ii + 1) * (ii + 2) / 2 <= k) ++ii;
      i = ii;
      j = k - (i * n - i * (i + 1) / 2) + i + 1;
      pred = pair_passes(M, S, i, j);
    }
    append_pair(pred, i, j, S, np);
  }
  wave_sync();
  return np;
}

// Segment position of global shape index s in world w's segment (-1 if it is not in it).
__device__ __forceinline__ int seg_position(const NaModel& M, int w, int nl, int s) {
  int s0 = M.world_shape_start[w];
  if (s >= s0 && s < s0 + nl) return s - s0;
  for (int k = 0; k < M.G; ++k)
    if (M.global_shapes[k] == s) return nl + k;
  return -1;
}

// Stage 2c: explicit broad phase (broad_phase_nxn.py:394-470): the world's precomputed pairs
// (already filtered by world, group and exclusion rules) tested for AABB overlap, in list order.
__device__ int stage_broad_explicit(const NaModel& M, int w, int nl, SegLDS& S) {
  const int p0 = M.world_pair_start[w], p1 = M.world_pair_start[w + 1];
  int np = 0;
  for (int base = p0; base < p1; base += NA_WAVE) {
    int k = base + lane_id();
    bool pred = false;
    int i = 0, j = 0;
    if (k < p1) {
      i = seg_position(M, w, nl, M.pair_list[2 * k]);
      j = seg_position(M, w, nl, M.pair_list[2 * k + 1]);
      if (i > j) { int t = i; i = j; j = t; }
      pred = i >= 0 && S.lo[0][i] <= S.hi[0][j] && S.hi[0][i] >= S.lo[0][j] && S.lo[1][i] <= S.hi[1][j] &&
             S.hi[1][i] >= S.lo[1][j] && S.lo[2][i] <= S.hi[2][j] && S.hi[2][i] >= S.lo[2][j];
    }
    append_pair(pred, i, j, S, np);
  }
  wave_sync();
  return np;
}

// Candidate pairs of world w's segment into S.pairs (0 sweep and prune, 1 all pairs, 2 explicit).
__device__ __forceinline__ int stage_broad(const NaModel& M, int w, int nseg, int nl, int broad_phase, SegLDS& S);

// Stage 2b: sweep-and-prune in LDS (broad_phase_sap.py:144-394): project on the fixed axis,
// bitonic sort of the lower bounds, binary-search ranges, scan, sweep.
__device__ int stage_broad_sap(const NaModel& M, int n, SegLDS& S) {
  if (n < 2) return 0;
  const float dx = 0.5935f, dy = 0.7790f, dz = 0.1235f;
  const float il = 1.f / sqrtf(dx * dx + dy * dy + dz * dz);
  const float ax = dx * il, ay = dy * il, az = dz * il;
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = lane_id(); i < P; i += NA_WAVE) {
    if (i < n) {
      float cx = 0.5f * (S.lo[0][i] + S.hi[0][i]), cy = 0.5f * (S.lo[1][i] + S.hi[1][i]),
            cz = 0.5f * (S.lo[2][i] + S.hi[2][i]);
      float hx = 0.5f * (S.hi[0][i] - S.lo[0][i]), hy = 0.5f * (S.hi[1][i] - S.lo[1][i]),
            hz = 0.5f * (S.hi[2][i] - S.lo[2][i]);
      float c = ax * cx + ay * cy + az * cz, r = ax * hx + ay * hy + az * hz;
      S.proj_lo[i] = c - r;
      S.proj_hi[i] = c + r;
      S.order[i] = i;
    } else {
      S.proj_lo[i] = 1e30f;
      S.proj_hi[i] = 1e30f;
      S.order[i] = -1;
    }
  }
  wave_sync();
  // bitonic sort of (proj_lo, order) pairs; keys of the sorted position stay in scan[]
  for (int i = lane_id(); i < P; i += NA_WAVE) S.scan[i] = __float_as_int(S.proj_lo[i]);
  wave_sync();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane_id(); i < P; i += NA_WAVE) {
        int ixj = i ^ j;
        if (ixj > i) {
          float a = __int_as_float(S.scan[i]), b = __int_as_float(S.scan[ixj]);
          int oa = S.order[i], ob = S.order[ixj];
          bool up = (i & k) == 0;
          // order by (key, original index) so ties sort deterministically
          bool gt = (a > b) || (a == b && (unsigned)oa > (unsigned)ob);
          if (gt == up) {
            S.scan[i] = __float_as_int(b); S.scan[ixj] = __float_as_int(a);
            S.order[i] = ob; S.order[ixj] = oa;
          }
        }
      }
      wave_sync();
    }
  }
  // ranges: number of later sorted elements whose lower bound is < this element's upper bound
  int my_range[NA_MAX_SEG / NA_WAVE];
#pragma unroll
  for (int t = 0; t < NA_MAX_SEG / NA_WAVE; ++t) {
    int i = lane_id() + t * NA_WAVE;
    my_range[t] = 0;
    if (i < n) {
      float up = S.proj_hi[S.order[i]];
      int lo = i + 1, hi = n;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (__int_as_float(S.scan[mid]) < up) lo = mid + 1;
        else hi = mid;
      }
      my_range[t] = lo - i - 1;
    }
  }
  wave_sync();
  // inclusive scan of ranges over sorted positions (chunks of 64)
  int carry = 0;
#pragma unroll
  for (int t = 0; t < NA_MAX_SEG / NA_WAVE; ++t) {
    int i = lane_id() + t * NA_WAVE;
    int incl = wave_incl_scan(my_range[t]) + carry;
    if (i < NA_MAX_SEG) S.scan[i] = incl;
    carry = __shfl(incl, NA_WAVE - 1, NA_WAVE);
  }
  wave_sync();
  int total = S.scan[n - 1];
  int np = 0;
  for (int base = 0; base < total; base += NA_WAVE) {
    int k = base + lane_id();
    bool pred = false;
    int a = 0, b = 0;
    if (k < total) {
      // sorted position i: first with scan[i] > k
      int lo = 0, hi = n;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (S.scan[mid] > k) hi = mid;
        else lo = mid + 1;
      }
      int i = lo;
      int j = i + 1 + (k - (i > 0 ? S.scan[i - 1] : 0));
      a = S.order[i];
      b = S.order[j];
      if (a > b) { int t = a; a = b; b = t; }
      pred = pair_passes(M, S, a, b);
    }
    append_pair(pred, a, b, S, np);
  }
  wave_sync();
  return np;
}

// Narrow phase of one pair: manifold, radii, normal order (narrow_phase.py:154-604).
// Returns the number of valid (gap-checked) contacts written into out arrays.
struct PairOut {
  int mask;    // bit k set when manifold slot k passed the gap check
  int sa, sb;  // segment positions after type ordering
  Manifold m;
  float phi0[NA_MF];
  float ra, rb;
};
__device__ __forceinline__ void narrow_pair(const NaModel& M, const SegLDS& S, int i, int j, PairOut& o) {
  int ga = S.gidx[i], gb = S.gidx[j];
  int ta = M.shape_type[ga], tb = M.shape_type[gb];
  if (ta > tb) {
    int t = i; i = j; j = t;
    t = ga; ga = gb; gb = t;
    t = ta; ta = tb; tb = t;
  }
  o.sa = i;
  o.sb = j;
  o.mask = 0;
  f3 pa = mk(S.pose[0][i], S.pose[1][i], S.pose[2][i]);
  f3 pb = mk(S.pose[0][j], S.pose[1][j], S.pose[2][j]);
  m3 Ra = qmat(S.pose[3][i], S.pose[4][i], S.pose[5][i], S.pose[6][i]);
  m3 Rb = qmat(S.pose[3][j], S.pose[4][j], S.pose[5][j], S.pose[6][j]);
  const float* da = M.shape_scale + 3 * ga;
  const float* db = M.shape_scale + 3 * gb;
  f3 sca = mk(da[0], da[1], da[2]), scb = mk(db[0], db[1], db[2]);
  float gap_sum = M.shape_gap[ga] + M.shape_gap[gb];
  float ma = M.shape_margin[ga], mb = M.shape_margin[gb];
  float ra = (ta == GEO_SPHERE || ta == GEO_CAPSULE) ? sca.x : 0.f;
  float rb = (tb == GEO_SPHERE || tb == GEO_CAPSULE) ? scb.x : 0.f;
  o.ra = ra;
  o.rb = rb;
  Manifold& m = o.m;
  mf_clear(m);
  f3 za = col3(Ra, 2), zb = col3(Rb, 2);
  if (ta == GEO_PLANE && tb == GEO_SPHERE) {
    plane_sphere(za, pa, pb, scb.x, m.dist[0], m.pos[0]);
    m.nrm[0] = za;
  } else if (ta == GEO_PLANE && tb == GEO_BOX) {
    plane_box(za, pa, pb, Rb, scb, gap_sum, m);
  } else if (ta == GEO_SPHERE && tb == GEO_SPHERE) {
    sphere_sphere(pa, sca.x, pb, scb.x, m.dist[0], m.pos[0], m.nrm[0]);
  } else if (ta == GEO_PLANE && tb == GEO_CAPSULE) {
    f3 seg = zb * scb.y;
    plane_sphere(za, pa, pb + seg, scb.x, m.dist[0], m.pos[0]);
    plane_sphere(za, pa, pb - seg, scb.x, m.dist[1], m.pos[1]);
    m.nrm[0] = za;
    m.nrm[1] = za;
  } else if (ta == GEO_SPHERE && tb == GEO_CAPSULE) {
    sphere_capsule(pa, sca.x, pb, zb, scb.x, scb.y, m.dist[0], m.pos[0], m.nrm[0]);
  } else if (ta == GEO_CAPSULE && tb == GEO_CAPSULE) {
    capsule_capsule(pa, za, sca.x, sca.y, pb, zb, scb.x, scb.y, m);
  } else if (ta == GEO_SPHERE && tb == GEO_BOX) {
    sphere_box(pa, sca.x, pb, Rb, scb, m.dist[0], m.pos[0], m.nrm[0]);
  } else if (ta == GEO_CAPSULE && tb == GEO_BOX) {
    capsule_box(pa, za, sca.x, sca.y, pb, Rb, scb, m);
  } else if (ta == GEO_BOX && tb == GEO_BOX) {
    box_box(pa, Ra, sca, pb, Rb, scb, gap_sum + ma + mb, m);
  } else {
    return;
  }
#pragma unroll
  for (int k = 0; k < NA_MF; ++k) {
    if (!(m.dist[k] < NA_MAXVAL)) continue;
    f3 n = nrm3(m.nrm[k]);
    f3 qa = m.pos[k] - n * (0.5f * m.dist[k] + ra);
    f3 qb = m.pos[k] + n * (0.5f * m.dist[k] + rb);
    float d = dot3(qb - qa, n) - (ra + rb + ma + mb);
    if (!(d <= gap_sum)) continue;
    m.nrm[k] = n;
    o.phi0[k] = d;
    o.mask |= 1 << k;
  }
}

// MuJoCo row parameters of each contact of a set (contact law 1), mixed per pair at contact
// generation: solref time constant and damping ratio, dmax, and the impedance at the contact's
// position (ContactSet::k holds the impedance and ContactSet::mu the friction, 0 for condim 1).
template <int N>
struct MjContactParams {
  float tc[N], dr[N], dmax[N];
};
#define NA_MJ_MINVAL 1.0e-15f
#define NA_MJ_MINIMP 0.0001f
#define NA_MJ_MAXIMP 0.9999f
#define NA_MJ_MINMU 1.0e-5f

// Impedance d(pos) of solimp (dmin, dmax, width, midpoint, power): the two-piece power curve from
// dmin at |pos| = 0 to dmax at |pos| >= width (MuJoCo's solimp definition).
__device__ __forceinline__ float mj_impedance(const float* si, float pos) {
  float dmin = fminf(fmaxf(si[0], NA_MJ_MINIMP), NA_MJ_MAXIMP), dmax = fminf(fmaxf(si[1], NA_MJ_MINIMP), NA_MJ_MAXIMP);
  float width = fmaxf(NA_MJ_MINVAL, si[2]), mid = fminf(fmaxf(si[3], NA_MJ_MINIMP), NA_MJ_MAXIMP);
  float power = fmaxf(1.f, si[4]);
  float x = fabsf(pos) / width;
  if (x >= 1.f) return dmax;
  float y;
  if (power == 1.f) y = x;
  else if (x < mid) y = powf(x, power) / powf(mid, power - 1.f);
  else y = 1.f - powf(1.f - x, power) / powf(1.f - mid, power - 1.f);
  return fminf(fmaxf(dmin + y * (dmax - dmin), dmin), dmax);
}

// Contact parameters of a shape pair (contact_params, solvers/mujoco/kernels.py:127-190): the
// higher priority shape decides; equal priorities mix solref/solimp by solmix (0.5 when both are
// ~0, one-sided when one is), take the larger condim and the larger friction; a non-positive
// solref time constant on either side selects the elementwise min.  Friction floored at MJ_MINMU.
struct MjPair {
  float solref[2], solimp[5], mu;
  int condim;
};
__device__ __forceinline__ MjPair mj_pair_params(const NaModel& M, int a, int b) {
  MjPair r;
  int pa = M.mj_priority[a], pb = M.mj_priority[b];
  float mix;
  if (pa != pb) {
    const int s = pa > pb ? a : b;
    mix = pa > pb ? 1.f : 0.f;
    r.condim = M.mj_condim[s];
    r.mu = M.shape_mu[s];
  } else {
    float s1 = M.mj_solmix[a], s2 = M.mj_solmix[b];
    const float tiny = 2.220446049250313e-16f;
    mix = s1 + s2 != 0.f ? s1 / (s1 + s2) : 0.f;
    if (s1 < tiny && s2 < tiny) mix = 0.5f;
    else if (s1 < tiny) mix = 0.f;
    else if (s2 < tiny) mix = 1.f;
    r.condim = max(M.mj_condim[a], M.mj_condim[b]);
    r.mu = fmaxf(M.shape_mu[a], M.shape_mu[b]);
  }
  r.mu = fmaxf(NA_MJ_MINMU, r.mu);
  const float *ra = M.mj_solref + 2 * a, *rb = M.mj_solref + 2 * b;
  const bool both = ra[0] > 0.f && rb[0] > 0.f;
#pragma unroll
  for (int k = 0; k < 2; ++k) r.solref[k] = both ? mix * ra[k] + (1.f - mix) * rb[k] : fminf(ra[k], rb[k]);
#pragma unroll
  for (int k = 0; k < 5; ++k) r.solimp[k] = mix * M.mj_solimp[5 * a + k] + (1.f - mix) * M.mj_solimp[5 * b + k];
  return r;
}

// Selection key of a contact when a world generates more contacts than its set holds: the gap
// distance phi0 (as an order-preserving integer), then the pair's segment positions (lower,
// higher) and the manifold slot.  Keys are unique within a world, so "the cap smallest keys" is
// one set whatever order the contacts come in.
// Position of the n-th set bit (n from 0) of m; n must be below popcount(m).
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int n) {
  int p = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int c = __popcll(m & ((1ull << w) - 1ull));
    if (n >= c) {
      n -= c;
      m >>= w;
      p += w;
    }
  }
  return p;
}
__device__ __forceinline__ unsigned long long contact_key(float phi0, int sa, int sb, int sub) {
  unsigned int b = __float_as_uint(phi0);
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  const unsigned int lo = (unsigned int)min(sa, sb), hi = (unsigned int)max(sa, sb);
  return ((unsigned long long)b << 32) | (lo << 16) | (hi << 8) | (unsigned int)sub;
}

// Stage 3: narrow phase over the pair list, compacted into the LDS contact set in pair order.
// Returns the number of contacts generated.  Up to `cap` (which must not exceed the capacity of
// the contact set) they are all kept in pair order; beyond it the set keeps the `cap` contacts
// with the smallest gap distance (contact_key): once the set is full, each pass of 64 pairs keeps
// the `cap` smallest keys of the set and the pass's contacts, its kept contacts taking the freed
// slots.  The reference appends every contact to one pool of 128 per world
// (rigid_contact_max, solver_sap_adaptive.py:573-577) and drops what overflows the pool; a world
// here never holds more than its LDS set, and what it drops is the farthest contacts.  With `mj`
// (MuJoCo contact law) each contact also gets its MuJoCo row parameters, and k / mu hold its
// impedance and friction.
template <class CS, class MJ = MjContactParams<1>>
__device__ int stage_narrow(const NaModel& M, int b0, int np, int cap, const SegLDS& S, CS& C, MJ* mj = nullptr) {
  int nc = 0;
  np = min(np, NA_MAX_PAIRS);
  for (int base = 0; base < np; base += NA_WAVE) {
    int k = base + lane_id();
    PairOut o;
    o.mask = 0;
    if (k < np) {
      uint32_t pr = S.pairs[k];
      narrow_pair(M, S, (int)(pr >> 8), (int)(pr & 0xff), o);
    }
    // this lane's contact of manifold slot q into set slot `slot`
    auto put = [&](int slot, int q, int ba, int bb, float kk, float mu, const MjPair& mp) {
      C.ba[slot] = (int8_t)(ba < 0 ? -1 : ba - b0);
      C.bb[slot] = (int8_t)(bb < 0 ? -1 : bb - b0);
      C.sa[slot] = (uint8_t)o.sa;
      C.sb[slot] = (uint8_t)o.sb;
      C.sub[slot] = (uint8_t)q;
      C.n[0][slot] = o.m.nrm[q].x; C.n[1][slot] = o.m.nrm[q].y; C.n[2][slot] = o.m.nrm[q].z;
      C.p[0][slot] = o.m.pos[q].x; C.p[1][slot] = o.m.pos[q].y; C.p[2][slot] = o.m.pos[q].z;
      C.dist[slot] = o.m.dist[q];
      C.phi0[slot] = o.phi0[q];
      C.k[slot] = kk;
      C.mu[slot] = mu;
      C.ra[slot] = o.ra;
      C.rb[slot] = o.rb;
      if (mj) {
        // the row position: dist - margins, or the surface distance where margins are zeroed
        const float pos = M.mj_zero_margins ? o.m.dist[q] : o.phi0[q];
        C.k[slot] = mj_impedance(mp.solimp, pos);
        mj->tc[slot] = mp.solref[0];
        mj->dr[slot] = mp.solref[1];
        mj->dmax[slot] = fminf(fmaxf(mp.solimp[1], NA_MJ_MINIMP), NA_MJ_MAXIMP);
      }
    };
    int cnt = __popc(o.mask);
    int incl = wave_incl_scan(cnt);
    int excl = incl - cnt;
    int tot = __shfl(incl, NA_WAVE - 1, NA_WAVE);
    if (cnt > 0) {
      int ga = S.gidx[o.sa], gb = S.gidx[o.sb];
      int ba = M.shape_body[ga], bb = M.shape_body[gb];
      float kk = 0.5f * (M.shape_ke[ga] + M.shape_ke[gb]);
      float mu = fmaxf(M.shape_mu[ga], M.shape_mu[gb]);
      MjPair mp;
      if (mj) {
        mp = mj_pair_params(M, ga, gb);
        mu = mp.condim == 1 ? 0.f : mp.mu;
      }
      int slot = nc + excl;
#pragma unroll
      for (int q = 0; q < NA_MF; ++q) {
        if (!((o.mask >> q) & 1)) continue;
        if (slot < cap) put(slot, q, ba, bb, kk, mu, mp);
        slot++;
      }
    }
    if (nc + tot > cap) {
      // the set is full: it keeps the cap smallest keys of (set, this pass's contacts past it).
      // T = the cap-th smallest of those keys (bisection with wave-wide counts);
      // set slots with a larger key are freed, in slot order, and the pass's contacts with a key
      // <= T fill them in generation order.  Keys are unique, so exactly cap keys are <= T.
      wave_sync();
      const int l = lane_id();
      const unsigned long long none = ~0ull;  // above every contact's key
      unsigned long long nk[NA_MF];
      int gen = nc + excl;  // this lane's next generation index
#pragma unroll
      for (int q = 0; q < NA_MF; ++q) {
        const bool has = (o.mask >> q) & 1;
        nk[q] = (has && gen >= cap) ? contact_key(o.phi0[q], o.sa, o.sb, q) : none;
        gen += has ? 1 : 0;
      }
      const unsigned long long sk0 = l < cap ? contact_key(C.phi0[l], C.sa[l], C.sb[l], C.sub[l]) : none;
      const unsigned long long sk1 =
          l + NA_WAVE < cap ? contact_key(C.phi0[l + NA_WAVE], C.sa[l + NA_WAVE], C.sb[l + NA_WAVE], C.sub[l + NA_WAVE])
                            : none;
      // the distance bits (high word) first; the low word only where distances tie at T
      unsigned int lo = 0u, hi = 0xffffffffu;  // count(key >> 32 <= hi) >= cap
      while (lo < hi) {
        const unsigned int mid = lo + ((hi - lo) >> 1);
        int below = ((unsigned int)(sk0 >> 32) <= mid) + ((unsigned int)(sk1 >> 32) <= mid);
#pragma unroll
        for (int q = 0; q < NA_MF; ++q) below += (unsigned int)(nk[q] >> 32) <= mid;
        if (wave_sum_i(below) >= cap) hi = mid;
        else lo = mid + 1u;
      }
      const unsigned long long th = (unsigned long long)lo << 32;
      int le = (sk0 <= (th | 0xffffffffull)) + (sk1 <= (th | 0xffffffffull));
#pragma unroll
      for (int q = 0; q < NA_MF; ++q) le += nk[q] <= (th | 0xffffffffull);
      unsigned int tl = 0xffffffffu;
      if (wave_sum_i(le) > cap) {  // several keys share the distance bits at the threshold
        unsigned int a = 0u, b = 0xffffffffu;
        while (a < b) {
          const unsigned int mid = a + ((b - a) >> 1);
          int below = (sk0 <= (th | mid)) + (sk1 <= (th | mid));
#pragma unroll
          for (int q = 0; q < NA_MF; ++q) below += nk[q] <= (th | mid);
          if (wave_sum_i(below) >= cap) b = mid;
          else a = mid + 1u;
        }
        tl = a;
      }
      const unsigned long long T = th | tl;
      const unsigned long long f0 = __ballot(sk0 != none && sk0 > T);  // freed slots 0..63
      const unsigned long long f1 = __ballot(sk1 != none && sk1 > T);  // freed slots 64..127
      const int n0 = __popcll(f0);
      int mine = 0;
#pragma unroll
      for (int q = 0; q < NA_MF; ++q) mine += nk[q] <= T;
      int rank = wave_incl_scan(mine) - mine;
      if (mine > 0) {
        const int ga = S.gidx[o.sa], gb = S.gidx[o.sb];
        const int ba = M.shape_body[ga], bb = M.shape_body[gb];
        float mu = fmaxf(M.shape_mu[ga], M.shape_mu[gb]);
        MjPair mp;
        if (mj) {
          mp = mj_pair_params(M, ga, gb);
          mu = mp.condim == 1 ? 0.f : mp.mu;
        }
        const float kk = 0.5f * (M.shape_ke[ga] + M.shape_ke[gb]);
#pragma unroll
        for (int q = 0; q < NA_MF; ++q) {
          if (!(nk[q] <= T)) continue;
          const int slot = rank < n0 ? nth_set_bit(f0, rank) : NA_WAVE + nth_set_bit(f1, rank - n0);
          put(slot, q, ba, bb, kk, mu, mp);
          ++rank;
        }
      }
      wave_sync();
    }
    nc += tot;
  }
  wave_sync();
  return nc;
}


__device__ __forceinline__ int stage_broad(const NaModel& M, int w, int nseg, int nl, int broad_phase, SegLDS& S) {
  if (broad_phase == 2) return stage_broad_explicit(M, w, nl, S);
  return broad_phase == 1 ? stage_broad_nxn(M, nseg, S) : stage_broad_sap(M, nseg, S);
}

// Contact set in LDS -> HBM for the parity entry points: per contact NA_DUMP_FLOATS floats
// (shape a, shape b, body a, body b (global, -1 static), manifold slot, normal A->B, midpoint,
// distance, phi0, stiffness, friction, 0), exactly the values the convex solve reads.  Must run
// before the solve set-up, which reuses the segment's LDS.
#define NA_DUMP_FLOATS 16
template <class CS>
__device__ void dump_contacts(const SegLDS& seg, const CS& con, int b0, int nc, float* out) {
  for (int c = lane_id(); c < nc; c += NA_WAVE) {
    float* o = out + (long long)c * NA_DUMP_FLOATS;
    o[0] = (float)seg.gidx[con.sa[c]];
    o[1] = (float)seg.gidx[con.sb[c]];
    o[2] = (float)(con.ba[c] >= 0 ? b0 + con.ba[c] : -1);
    o[3] = (float)(con.bb[c] >= 0 ? b0 + con.bb[c] : -1);
    o[4] = (float)con.sub[c];
#pragma unroll
    for (int k = 0; k < 3; ++k) { o[5 + k] = con.n[k][c]; o[8 + k] = con.p[k][c]; }
    o[11] = con.dist[c];
    o[12] = con.phi0[c];
    o[13] = con.k[c];
    o[14] = con.mu[c];
    o[15] = 0.f;
  }
}
