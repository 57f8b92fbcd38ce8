// Iterate / wave storage types.  Arithmetic is always fp32; only HBM storage narrows.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace qpch::dev {

// Raw buffer resources: one start's array addressed by 32-bit byte offsets, a lane's own
// offset in a VGPR plus a wave-uniform one in an SGPR, so that a strided sequence of loads
// costs no per-lane address arithmetic.  Word 3 (0x00020000, 32-bit data format) is the
// gfx9 family's raw-buffer setting.  Callers keep every offset inside the array (the
// launcher guarantees arrays below 2 GiB per start).
using Rsrc = __amdgpu_buffer_rsrc_t;

// Loads that another CU of the same XCD may just have written (the persistent launches):
// device scope, which does not hit in the CU's L1 (a plain or work-group-scope load, or a
// work-group-scope L1 invalidate, returned the CU's stale L1 copy every time in a probe on
// the MI355X; a device-scope load never did).  kCoh is the buffer instructions' cache
// policy bit for it (SC1); pointer loads get it as relaxed agent-scope atomic loads.
constexpr int kCoh = 16;
template <class T>
__device__ __forceinline__ T load_coherent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ Rsrc make_rsrc(const void* base, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}
__device__ __forceinline__ float buf_f32(Rsrc r, int vo, int so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
__device__ __forceinline__ unsigned char buf_u8(Rsrc r, int vo, int so) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, vo, so, 0);
}

template <class St>
struct Storage;

template <>
struct Storage<float2> {
  template <bool COH = false>
  static __device__ __forceinline__ float2 ld(const float2* p, size_t i) {
    if constexpr (COH) {
      return __builtin_bit_cast(float2, load_coherent(reinterpret_cast<const unsigned long long*>(p + i)));
    } else {
      return p[i];
    }
  }
  static __device__ __forceinline__ void st(float2* p, size_t i, float2 v) { p[i] = v; }
  // byte offsets vo (lane) + so (uniform) into a buffer resource
  template <bool COH = false>
  static __device__ __forceinline__ float2 bld(Rsrc r, int vo, int so) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, COH ? kCoh : 0));
  }
  static __device__ __forceinline__ void bst(Rsrc r, int vo, int so, float2 v) {
    using U = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, vo, so, 0);
  }
};

template <>
struct Storage<__half2> {
  template <bool COH = false>
  static __device__ __forceinline__ float2 ld(const __half2* p, size_t i) {
    if constexpr (COH) {
      return __half22float2(__builtin_bit_cast(__half2, load_coherent(reinterpret_cast<const unsigned*>(p + i))));
    } else {
      return __half22float2(p[i]);
    }
  }
  static __device__ __forceinline__ void st(__half2* p, size_t i, float2 v) {
    p[i] = __float22half2_rn(v);
  }
  template <bool COH = false>
  static __device__ __forceinline__ float2 bld(Rsrc r, int vo, int so) {
    return __half22float2(__builtin_bit_cast(__half2, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, COH ? kCoh : 0)));
  }
  static __device__ __forceinline__ void bst(Rsrc r, int vo, int so, float2 v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, __float22half2_rn(v)), r, vo, so, 0);
  }
};

template <>
struct Storage<__hip_bfloat162> {
  template <bool COH = false>
  static __device__ __forceinline__ float2 ld(const __hip_bfloat162* p, size_t i) {
    if constexpr (COH) {
      const unsigned u = load_coherent(reinterpret_cast<const unsigned*>(p + i));
      __hip_bfloat162 h;
      __builtin_memcpy(static_cast<void*>(&h), &u, sizeof(u));
      return __bfloat1622float2(h);
    } else {
      return __bfloat1622float2(p[i]);
    }
  }
  static __device__ __forceinline__ void st(__hip_bfloat162* p, size_t i, float2 v) {
    p[i] = __float22bfloat162_rn(v);
  }
  // (__hip_bfloat162 is not trivially copyable: its bits move through memcpy)
  template <bool COH = false>
  static __device__ __forceinline__ float2 bld(Rsrc r, int vo, int so) {
    const unsigned u = __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, COH ? kCoh : 0);
    __hip_bfloat162 h;
    __builtin_memcpy(static_cast<void*>(&h), &u, sizeof(u));
    return __bfloat1622float2(h);
  }
  static __device__ __forceinline__ void bst(Rsrc r, int vo, int so, float2 v) {
    const __hip_bfloat162 h = __float22bfloat162_rn(v);
    unsigned u;
    __builtin_memcpy(&u, &h, sizeof(u));
    __builtin_amdgcn_raw_buffer_store_b32(u, r, vo, so, 0);
  }
};

}  // namespace qpch::dev
