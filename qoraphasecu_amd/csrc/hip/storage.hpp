// Iterate / wave storage types.  Arithmetic is always fp32; only HBM storage narrows.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace qpch::dev {

template <class St>
struct Storage;

template <>
struct Storage<float2> {
  static __device__ __forceinline__ float2 ld(const float2* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(float2* p, size_t i, float2 v) { p[i] = v; }
};

template <>
struct Storage<__half2> {
  static __device__ __forceinline__ float2 ld(const __half2* p, size_t i) {
    return __half22float2(p[i]);
  }
  static __device__ __forceinline__ void st(__half2* p, size_t i, float2 v) {
    p[i] = __float22half2_rn(v);
  }
};

template <>
struct Storage<__hip_bfloat162> {
  static __device__ __forceinline__ float2 ld(const __hip_bfloat162* p, size_t i) {
    return __bfloat1622float2(p[i]);
  }
  static __device__ __forceinline__ void st(__hip_bfloat162* p, size_t i, float2 v) {
    p[i] = __float22bfloat162_rn(v);
  }
};

}  // namespace qpch::dev
