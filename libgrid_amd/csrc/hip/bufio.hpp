// Raw buffer loads / stores: a 32-bit per-thread byte offset (VGPR) plus a uniform byte
// offset (SGPR), with the hardware range check of the buffer resource (loads past the end
// return 0, stores past the end are dropped).
#pragma once
#include <hip/hip_runtime.h>

namespace lg {

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned long bytes) {
  const int num = bytes >= 0xffffffffUL ? (int)0xffffffff : (int)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, num, 0x00020000);
}
template <typename C> __device__ __forceinline__ C bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <> __device__ __forceinline__ double2 bload<double2>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  double2 d;
  __builtin_memcpy(&d, &x, 16);
  return d;
}
template <> __device__ __forceinline__ float2 bload<float2>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  float2 d;
  __builtin_memcpy(&d, &x, 8);
  return d;
}
__device__ __forceinline__ double bload_f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  double d;
  __builtin_memcpy(&d, &x, 8);
  return d;
}
template <typename C> __device__ __forceinline__ void bstore(C v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <> __device__ __forceinline__ void bstore<double2>(double2 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  __attribute__((ext_vector_type(4))) unsigned x;
  __builtin_memcpy(&x, &v, 16);
  __builtin_amdgcn_raw_buffer_store_b128(x, r, voff, soff, 0);
}
template <> __device__ __forceinline__ void bstore<float2>(float2 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  __attribute__((ext_vector_type(2))) unsigned x;
  __builtin_memcpy(&x, &v, 8);
  __builtin_amdgcn_raw_buffer_store_b64(x, r, voff, soff, 0);
}
__device__ __forceinline__ void bstore_f64(double v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  __attribute__((ext_vector_type(2))) unsigned x;
  __builtin_memcpy(&x, &v, 8);
  __builtin_amdgcn_raw_buffer_store_b64(x, r, voff, soff, 0);
}

__device__ __forceinline__ float bload_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  const auto x = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  float d;
  __builtin_memcpy(&d, &x, 4);
  return d;
}

}  // namespace lg
