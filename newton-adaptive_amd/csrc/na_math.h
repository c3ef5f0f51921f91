// Device math for the CDNA4 step-doubling rigid-body kernels: float3 / quaternion helpers
// and 64-lane wavefront reductions and scans.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NA_WAVE 64

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 nrm3(f3 a) {
  float l = len3(a);
  return l > 0.f ? a * (1.f / l) : a;
}
__device__ __forceinline__ float comp3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ void setc3(f3& a, int i, float s) {
  if (i == 0) a.x = s;
  else if (i == 1) a.y = s;
  else a.z = s;
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// rotation matrix (row-major) of a unit quaternion (x, y, z, w)
struct m3 {
  float m[9];
};
__device__ __forceinline__ m3 qmat(float x, float y, float z, float w) {
  m3 r;
  r.m[0] = 1.f - 2.f * (y * y + z * z);
  r.m[1] = 2.f * (x * y - z * w);
  r.m[2] = 2.f * (x * z + y * w);
  r.m[3] = 2.f * (x * y + z * w);
  r.m[4] = 1.f - 2.f * (x * x + z * z);
  r.m[5] = 2.f * (y * z - x * w);
  r.m[6] = 2.f * (x * z - y * w);
  r.m[7] = 2.f * (y * z + x * w);
  r.m[8] = 1.f - 2.f * (x * x + y * y);
  return r;
}
__device__ __forceinline__ f3 mv(const m3& a, f3 v) {
  return f3{a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z};
}
__device__ __forceinline__ f3 mtv(const m3& a, f3 v) {
  return f3{a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z, a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
            a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z};
}
__device__ __forceinline__ f3 col3(const m3& a, int j) { return f3{a.m[j], a.m[3 + j], a.m[6 + j]}; }

struct qf {
  float x, y, z, w;
};
__device__ __forceinline__ qf qmul(qf a, qf b) {
  return qf{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__device__ __forceinline__ f3 qrot(qf q, f3 v) { return mv(qmat(q.x, q.y, q.z, q.w), v); }

// ---------------------------------------------------------------- wave collectives
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, NA_WAVE);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, NA_WAVE);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, NA_WAVE));
  return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o, NA_WAVE);
    v = (u > v || u != u) ? u : v;  // NaN propagates
  }
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, NA_WAVE);
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, NA_WAVE));
  return v;
}
// inclusive prefix sum over the 64 lanes (Hillis-Steele with shuffles)
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < NA_WAVE; o <<= 1) {
    int u = __shfl_up(v, o, NA_WAVE);
    if (l >= o) v += u;
  }
  return v;
}
// rank of this lane among lanes whose predicate is set (ballot + mbcnt compaction)
__device__ __forceinline__ int ballot_rank(bool pred, int* total) {
  unsigned long long m = __ballot(pred);
  *total = __popcll(m);
  unsigned long long below = m & ((1ull << lane_id()) - 1ull);
  return __popcll(below);
}

// Keep LDS traffic of one wave ordered around cross-lane exchange.  A workgroup is exactly
// one wavefront, so the barrier is cheap; it also orders LDS writes before reads.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }
