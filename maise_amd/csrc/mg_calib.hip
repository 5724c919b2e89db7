// mg_calib.hip -- issue-rate microbenchmark of the vector ALU operations the descriptor and force kernels are built
// from, measured on the device instead of assumed: FP32 FMA (non-packed: packed FP32 is switched off in this build)
// and the transcendental unit (exp2, cos, sqrt, rcp).  The rates price the roofline of those kernels.
//
// Each lane runs 8 independent dependency chains of one operation (enough to cover the operation's latency with
// the waves of a fully occupied SIMD), ITERS steps each, so the loop is bound by the issue rate of that operation.
#include <hip/hip_runtime.h>

#include "mg_internal.h"

#define CAL_ITERS 4096
#define CAL_THREADS 256
#define CAL_BLOCKS 8192  // 32 waves per CU on 256 CUs: every SIMD fully occupied

template <int OP>
__global__ __launch_bounds__(CAL_THREADS) void k_calib(float* __restrict__ out, float a, float b) {
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = 0.25f + 0.01f * (float)((threadIdx.x + k) & 15);
  for (int it = 0; it < CAL_ITERS; it++) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (OP == 0) x[k] = __builtin_fmaf(x[k], a, b);          // v_fma_f32
      if (OP == 1) x[k] = __builtin_amdgcn_exp2f(-x[k]);       // v_exp_f32 (stays in (0, 1])
      if (OP == 2) x[k] = __builtin_amdgcn_cosf(x[k]);         // v_cos_f32 (revolutions; stays in [-1, 1])
      if (OP == 3) x[k] = __builtin_amdgcn_sqrtf(x[k] + 0.5f);  // v_sqrt_f32 (with one add)
      if (OP == 4) x[k] = __builtin_amdgcn_rcpf(x[k]);         // v_rcp_f32 (alternates x, 1/x)
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; k++) s += x[k];
  out[blockIdx.x * CAL_THREADS + threadIdx.x] = s;
}

extern "C" int mgk_status(void);

// Operations per second of op (0 FMA, 1 exp2, 2 cos, 3 sqrt, 4 rcp) on device, one wave lane per operation counted
// once: the mean of three timed launches after a warm-up.  The sqrt loop also issues one add per sqrt.  Returns 0 or
// -1 (the HIP error is kept for mgk_last_error()).
extern "C" int mg_calibrate_rate(int device, int op, double* ops_per_s) {
  if (op < 0 || op > 4 || !ops_per_s) return -1;
  if (hipSetDevice(device) != hipSuccess) return mgk_status();
  float* out = nullptr;
  hipEvent_t e0, e1;
  if (hipMalloc((void**)&out, sizeof(float) * CAL_BLOCKS * CAL_THREADS) != hipSuccess) return mgk_status();
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipFree(out);
    return mgk_status();
  }
  auto launch = [&]() {
    const float a = 0.999f, b = 0.001f;
    switch (op) {
      case 0: hipLaunchKernelGGL(k_calib<0>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, 0, out, a, b); break;
      case 1: hipLaunchKernelGGL(k_calib<1>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, 0, out, a, b); break;
      case 2: hipLaunchKernelGGL(k_calib<2>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, 0, out, a, b); break;
      case 3: hipLaunchKernelGGL(k_calib<3>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, 0, out, a, b); break;
      default: hipLaunchKernelGGL(k_calib<4>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, 0, out, a, b); break;
    }
  };
  launch();
  int rc = 0;
  float ms = 0.f;
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < 3; r++) launch();
  (void)hipEventRecord(e1, 0);
  if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f) rc = -1;
  if (mgk_status() != 0) rc = -1;
  if (rc == 0) {
    const double ops = 3.0 * (double)CAL_BLOCKS * CAL_THREADS * CAL_ITERS * 8.0;
    *ops_per_s = ops / (1e-3 * (double)ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc == 0 ? 0 : -1;
}
