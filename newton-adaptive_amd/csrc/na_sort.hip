// Deterministic contact order: the 63-bit sort key of
// newton/_src/geometry/contact_data.py:59-83 (make_contact_sort_key) and the radix sort plus
// gather of newton/_src/geometry/contact_sort.py:194-420 (ContactSorter.sort_full), over the
// Contacts SoA that na_collide writes.
//
// Key layout (bit 63 zero):  [62:43] shape0 (20 bits) | [42:23] shape1 (20 bits) | [22:0] sub.
// `sub` is the rank of the contact among the contacts of its shape pair in production order (the
// reference writes the manifold slot there; both keep the within-pair production order).  Slots at
// and past min(count, capacity) carry the sentinel 0x7FFFFFFFFFFFFFFF and keep their place.
//
// Everything runs on the caller's stream with device-side counts (no host synchronization), over
// the full capacity like the reference's sorter: three HBM-bound passes (keys, one radix sort of
// (key, index) pairs, one staged gather of every contact array).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "../../include/nadapt.h"

int na_set_error(const char* msg);  // nadapt.hip

namespace {

constexpr unsigned long long kSentinel = 0x7FFFFFFFFFFFFFFFull;
constexpr int kBlock = 256;
// staged contact words per slot: shape0, shape1, point0[3], point1[3], offset0[3], offset1[3],
// normal[3], margin0, margin1, tids, match
constexpr int kWords = 21;

inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

struct Layout {
  size_t keys_in, idx_in, perm, stage, cub, cub_bytes, total;
};

Layout layout(int cap, size_t cub_bytes) {
  Layout L;
  size_t o = 0;
  L.keys_in = o; o += align256(sizeof(unsigned long long) * size_t(cap));
  L.idx_in = o; o += align256(sizeof(int32_t) * size_t(cap));
  L.perm = o; o += align256(sizeof(int32_t) * size_t(cap));
  L.stage = o; o += align256(sizeof(int32_t) * kWords * size_t(cap));
  L.cub = o; o += align256(cub_bytes);
  L.cub_bytes = cub_bytes;
  L.total = o;
  return L;
}

int bit_length(long long v) {
  int b = 0;
  while (v > 0) { ++b; v >>= 1; }
  return b;
}

__global__ __launch_bounds__(kBlock) void k_sort_keys(const int32_t* count, int cap, const int32_t* s0,
                                                      const int32_t* s1, unsigned long long* keys,
                                                      int32_t* idx) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= cap) return;
  const int n = min(max(count[0], 0), cap);
  idx[i] = i;
  if (i >= n) { keys[i] = kSentinel; return; }
  const int a = s0[i], b = s1[i];
  // the contacts of one shape pair are written next to each other: count the run before i
  unsigned r = 0;
  for (int j = i - 1; j >= 0 && s0[j] == a && s1[j] == b && r < 0x7FFFFFu; --j) ++r;
  keys[i] = ((unsigned long long)(a & 0xFFFFF) << 43) | ((unsigned long long)(b & 0xFFFFF) << 23) |
            (unsigned long long)(r & 0x7FFFFFu);
}

struct Arrays {
  int32_t *s0, *s1;
  float *p0, *p1, *o0, *o1, *nrm, *m0, *m1;
  int32_t *tids, *match;
};

// copy the first n slots of every array into the stage (slot-major words)
__global__ __launch_bounds__(kBlock) void k_stage(const int32_t* count, int cap, Arrays A, int32_t* stage) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int n = min(max(count[0], 0), cap);
  if (i >= n) return;
  int32_t* w = stage + size_t(kWords) * i;
  w[0] = A.s0[i];
  w[1] = A.s1[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    w[2 + d] = __float_as_int(A.p0[3 * i + d]);
    w[5 + d] = __float_as_int(A.p1[3 * i + d]);
    w[8 + d] = __float_as_int(A.o0[3 * i + d]);
    w[11 + d] = __float_as_int(A.o1[3 * i + d]);
    w[14 + d] = __float_as_int(A.nrm[3 * i + d]);
  }
  w[17] = __float_as_int(A.m0[i]);
  w[18] = __float_as_int(A.m1[i]);
  w[19] = A.tids ? A.tids[i] : -1;
  w[20] = A.match ? A.match[i] : -1;
}

// slot i of the sorted order takes staged slot perm[i]; rank[perm[i]] = i; a match index k >= 0
// becomes match_remap[k] (the previous frame's sorted position of its contact k)
__global__ __launch_bounds__(kBlock) void k_gather(const int32_t* count, int cap, Arrays A, const int32_t* stage,
                                                   const int32_t* perm, const int32_t* match_remap,
                                                   int32_t remap_len, int32_t* rank) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int n = min(max(count[0], 0), cap);
  if (i >= cap) return;
  if (i >= n) {
    if (rank) rank[i] = i;
    return;
  }
  const int src = perm[i];  // < n: real keys sort before every sentinel
  const int32_t* w = stage + size_t(kWords) * src;
  if (rank) rank[src] = i;
  A.s0[i] = w[0];
  A.s1[i] = w[1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    A.p0[3 * i + d] = __int_as_float(w[2 + d]);
    A.p1[3 * i + d] = __int_as_float(w[5 + d]);
    A.o0[3 * i + d] = __int_as_float(w[8 + d]);
    A.o1[3 * i + d] = __int_as_float(w[11 + d]);
    A.nrm[3 * i + d] = __int_as_float(w[14 + d]);
  }
  A.m0[i] = __int_as_float(w[17]);
  A.m1[i] = __int_as_float(w[18]);
  if (A.tids) A.tids[i] = w[19];
  if (A.match) {
    int k = w[20];
    if (k >= 0 && match_remap) k = k < remap_len ? match_remap[k] : -1;
    A.match[i] = k;
  }
}

size_t cub_bytes_for(int cap) {
  size_t bytes = 0;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, bytes, static_cast<const unsigned long long*>(nullptr), static_cast<unsigned long long*>(nullptr),
      static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr), cap, 0, 63, hipStream_t(0));
  return e == hipSuccess ? bytes : size_t(0);
}

}  // namespace

extern "C" int64_t na_sort_contacts_workspace_bytes(int32_t capacity) {
  if (capacity < 1) return na_set_error("na_sort_contacts_workspace_bytes: capacity must be >= 1");
  const size_t cb = cub_bytes_for(capacity);
  if (cb == 0) return na_set_error("na_sort_contacts_workspace_bytes: radix sort size query failed");
  return int64_t(layout(capacity, cb).total);
}

extern "C" int na_sort_contacts(const na_contacts_desc* c, int32_t shape_count, int32_t* tids, int32_t* match_index,
                                const int32_t* match_remap, int32_t match_remap_len, int64_t* keys_out,
                                int32_t* rank_out, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!c || !c->count || !c->shape0 || !c->shape1 || !c->point0 || !c->point1 || !c->offset0 || !c->offset1 ||
      !c->normal || !c->margin0 || !c->margin1 || !keys_out || !workspace)
    return na_set_error("na_sort_contacts: null argument");
  const int cap = c->capacity;
  if (cap < 1) return na_set_error("na_sort_contacts: capacity must be >= 1");
  if (shape_count < 1 || shape_count > 0xFFFFF)
    return na_set_error("na_sort_contacts: shape_count must be in [1, 1048575] (20-bit key fields)");
  if (match_remap && match_remap_len < 0) return na_set_error("na_sort_contacts: negative match_remap_len");
  const size_t cb = cub_bytes_for(cap);
  if (cb == 0) return na_set_error("na_sort_contacts: radix sort size query failed");
  const Layout L = layout(cap, cb);
  if (workspace_bytes < int64_t(L.total)) return na_set_error("na_sort_contacts: workspace too small");
  char* ws = static_cast<char*>(workspace);
  auto* keys_in = reinterpret_cast<unsigned long long*>(ws + L.keys_in);
  auto* idx_in = reinterpret_cast<int32_t*>(ws + L.idx_in);
  auto* perm = reinterpret_cast<int32_t*>(ws + L.perm);
  auto* stage = reinterpret_cast<int32_t*>(ws + L.stage);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int grid = (cap + kBlock - 1) / kBlock;
  // shape indices are < shape_count, so the sentinel (all ones) stays last within the key bits
  const int end_bit = 43 + bit_length(shape_count);
  size_t cub_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, cub_bytes, static_cast<const unsigned long long*>(nullptr), static_cast<unsigned long long*>(nullptr),
      static_cast<const int32_t*>(nullptr), static_cast<int32_t*>(nullptr), cap, 0, end_bit, st);
  if (e != hipSuccess || cub_bytes > L.cub_bytes) return na_set_error("na_sort_contacts: radix sort size query failed");
  Arrays A{c->shape0, c->shape1, c->point0, c->point1, c->offset0, c->offset1, c->normal, c->margin0, c->margin1,
           tids, match_index};
  k_sort_keys<<<grid, kBlock, 0, st>>>(c->count, cap, c->shape0, c->shape1, keys_in, idx_in);
  k_stage<<<grid, kBlock, 0, st>>>(c->count, cap, A, stage);
  e = hipcub::DeviceRadixSort::SortPairs(ws + L.cub, cub_bytes, keys_in, reinterpret_cast<unsigned long long*>(keys_out),
                                         idx_in, perm, cap, 0, end_bit, st);
  if (e != hipSuccess) return na_set_error("na_sort_contacts: radix sort failed");
  k_gather<<<grid, kBlock, 0, st>>>(c->count, cap, A, stage, perm, match_remap, match_remap_len, rank_out);
  e = hipGetLastError();
  if (e != hipSuccess) return na_set_error(hipGetErrorString(e));
  return 0;
}
