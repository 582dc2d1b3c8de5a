// Raw buffer resources (stride 0) for gathers and stores whose masked or
// idle lanes must not branch: an access at an offset at or beyond
// num_records reads 0 and writes nothing, without touching memory. Device
// code only (included by the .hip kernel files).
#pragma once
#include <hip/hip_runtime.h>

namespace bpx {
namespace {

constexpr int kBufRange = 0x7FFFFFF8;         // num_records (bytes)
constexpr int kBufMasked = (int)0x80000000u;  // an offset beyond it
constexpr int kBufFlags = 0x00020000;         // dword 3: 32-bit data format (gfx9)

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const T* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, kBufRange, kBufFlags);
}

template <typename T>
__device__ inline T buffer_load(__amdgpu_buffer_rsrc_t r, int offset) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, offset, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, offset, 0, 0));
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ inline void buffer_store(T v, __amdgpu_buffer_rsrc_t r, int offset) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, offset, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, offset, 0, 0);
}

}  // namespace
}  // namespace bpx
