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

// Wave-wide reductions without LDS: DPP steps inside each 16-lane row (quad xor 1, quad
// xor 2, half-row mirror, row mirror) leave the row total in every lane of the row, then
// the four row totals are read from lanes 0/16/32/48 and combined in a fixed order.  Every
// lane returns the same bits.  Callers keep all 64 lanes active.
#define NA_DPP_XOR1 0xB1        // quad_perm [1,0,3,2]
#define NA_DPP_XOR2 0x4E        // quad_perm [2,3,0,1]
#define NA_DPP_HALF_MIRROR 0x141
#define NA_DPP_MIRROR 0x140

// bound_ctrl set: a lane whose source is invalid reads zero, so no zero-filled destination has to
// be prepared first (with every lane active, the row permutes below have no invalid source).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  long long b = __double_as_longlong(v);
  int lo = dpp_i<CTRL>((int)(b & 0xffffffffll));
  int hi = dpp_i<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_d(double v, int l) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Wave-uniform copies of values every lane holds alike (settings read through a pointer, loop
// bounds): the compiler then keeps them in scalar registers and branches on what depends on them
// without exec-mask bookkeeping.
__device__ __forceinline__ int uni_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uni_d(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// LDS-typed pointers for out-of-line functions that work on a world's LDS: a function that takes
// one and re-derives its references from it (lds_deref) compiles every access to a ds_*
// instruction with the field offset folded into the instruction.  Through a generic reference
// parameter the compiler casts each field's generic address to LDS at its use instead, with an
// address computation per field.
template <typename T>
using lds_ptr_t = __attribute__((address_space(3))) T*;
template <typename T>
__device__ __forceinline__ lds_ptr_t<T> lds_ptr(T& r) {
  return (lds_ptr_t<T>)(&r);
}
template <typename T>
__device__ __forceinline__ T& lds_deref(lds_ptr_t<T> p) {
  return *(T*)p;
}

struct OpAdd {
  template <typename T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMin {
  template <typename T>
  __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? a : b; }
};
struct OpMaxNan {  // max that propagates NaN from either side
  template <typename T>
  __device__ __forceinline__ T operator()(T a, T b) const { return (b > a || b != b) ? b : a; }
};

template <typename Op>
__device__ __forceinline__ float wave_reduce_f(float v, Op op) {
  v = op(v, dpp_f<NA_DPP_XOR1>(v));
  v = op(v, dpp_f<NA_DPP_XOR2>(v));
  v = op(v, dpp_f<NA_DPP_HALF_MIRROR>(v));
  v = op(v, dpp_f<NA_DPP_MIRROR>(v));
  return op(op(lane_f(v, 0), lane_f(v, 16)), op(lane_f(v, 32), lane_f(v, 48)));
}
template <typename Op>
__device__ __forceinline__ double wave_reduce_d(double v, Op op) {
  v = op(v, dpp_d<NA_DPP_XOR1>(v));
  v = op(v, dpp_d<NA_DPP_XOR2>(v));
  v = op(v, dpp_d<NA_DPP_HALF_MIRROR>(v));
  v = op(v, dpp_d<NA_DPP_MIRROR>(v));
  return op(op(lane_d(v, 0), lane_d(v, 16)), op(lane_d(v, 32), lane_d(v, 48)));
}
template <typename Op>
__device__ __forceinline__ int wave_reduce_i(int v, Op op) {
  v = op(v, dpp_i<NA_DPP_XOR1>(v));
  v = op(v, dpp_i<NA_DPP_XOR2>(v));
  v = op(v, dpp_i<NA_DPP_HALF_MIRROR>(v));
  v = op(v, dpp_i<NA_DPP_MIRROR>(v));
  return op(op(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
            op(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ __forceinline__ float wave_sum(float v) { return wave_reduce_f(v, OpAdd()); }
__device__ __forceinline__ double wave_sum_d(double v) { return wave_reduce_d(v, OpAdd()); }
__device__ __forceinline__ float wave_max(float v) { return wave_reduce_f(v, OpMaxNan()); }
__device__ __forceinline__ double wave_max_d(double v) { return wave_reduce_d(v, OpMaxNan()); }
__device__ __forceinline__ int wave_sum_i(int v) { return wave_reduce_i(v, OpAdd()); }
__device__ __forceinline__ int wave_min_i(int v) { return wave_reduce_i(v, OpMin()); }
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

// Order one wavefront's LDS traffic around a cross-lane exchange: the LDS operations of a wave
// execute in program order, so writes before the fence are seen by reads after it; the
// wavefront-scope fences keep the compiler from moving memory operations across.  It waits for
// no other wave, so it is also correct where one wave of a larger workgroup runs alone.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Barrier of every wave of the workgroup, with the LDS traffic before it visible after it.
__device__ __forceinline__ void group_sync() { __syncthreads(); }
