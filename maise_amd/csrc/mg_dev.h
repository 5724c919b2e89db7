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
  // tanh(x) = sign(x) (1 - e) / (1 + e), e = exp(-2|x|): no overflow, full FP32 accuracy
  float ax = fabsf(x);
  float e = __expf(-2.0f * ax);
  float t = (1.0f - e) / (1.0f + e);
  return copysignf(t, x);
}
