// Raw buffer resources for the MFMA kernels' operand loads: 32-bit per-lane
// byte offsets against a base held in scalar registers, with the hardware's
// range check (a load at or past num_records bytes returns zero, so tails
// need no clamping and masked lanes take an out-of-range offset). Callers
// keep every resource below 2^30 bytes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qmx {

constexpr int kRsrcWord3 = 0x00020000;   // fourth descriptor word, gfx9 family
constexpr int kOutOfRange = 0x7FFFFFF0;  // a byte offset past every resource here

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes) {
  const int n = bytes <= 0 ? 0 : (bytes < 0x7FFFFFFF ? (int)bytes : 0x7FFFFFFF);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, n, kRsrcWord3);
}

template <typename V>
__device__ __forceinline__ V rsrc_load(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ float2 rsrc_load<float2>(__amdgpu_buffer_rsrc_t r, int off) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  float2 v;
  __builtin_memcpy(&v, &x, sizeof(v));
  return v;
}
template <>
__device__ __forceinline__ double2 rsrc_load<double2>(__amdgpu_buffer_rsrc_t r, int off) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  double2 v;
  __builtin_memcpy(&v, &x, sizeof(v));
  return v;
}

// Four consecutive complex values from byte offset off, one load each, so
// that the range check falls on every value by itself.
template <typename C2>
__device__ __forceinline__ void rsrc_load4(__amdgpu_buffer_rsrc_t r, int off, C2 (&v)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = rsrc_load<C2>(r, off + (int)sizeof(C2) * e);
}

}  // namespace qmx
