// mg_dev.h -- device helpers shared by the gfx950 kernel files.
#pragma once
#include <hip/hip_runtime.h>

#define WAVE 64

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// LDS ordering between the lanes of one wave (several independent waves share a workgroup, so no block barrier)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float tanh_f(float x) {
  // tanh(x) = sign(x) (1 - e) / (1 + e), e = exp(-2|x|) = 2^(-2 log2(e) |x|) in (0, 1]: no overflow.  The hardware
  // exp2 and reciprocal (1 ulp each) in place of the IEEE division sequence: 7 instructions instead of 17, within a
  // few ulp of tanh (absolute error below 1.2e-7 where 1 - e cancels near 0)
  const float e = __builtin_amdgcn_exp2f(-2.0f * 1.4426950408889634f * __builtin_fabsf(x));
  const float t = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
  return __builtin_copysignf(t, x);
}
