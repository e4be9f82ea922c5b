// Shared device helpers and host-side error plumbing for libpyronn_mi.so.
//
// Arithmetic conventions (identical in every kernel of the library):
//  * The library is compiled with -ffp-contract=off; every fused multiply-add
//    is written out with fmaf().  Where the reference writes `a*b + c`, the
//    kernels use fmaf(a, b, c), the contraction nvcc applies by default.
//  * Division and square root are IEEE correctly rounded
//    (-fhip-fp32-correctly-rounded-divide-sqrt); rsqrtf(x) of the reference is
//    evaluated as 1.0f / sqrtf(x).
//  * The cone samplers clamp the coordinate into [-1, N] before the floor and
//    read neighbourhood cells built over index range [-1, N] with zeros outside
//    the data.  For every finite coordinate this is exactly the reference's
//    "zero outside [0, N-1]" rule without a bounds test per neighbour; NaN
//    coordinates read zeros.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "pyronn_mi.h"

namespace pyronn {

// ----------------------------------------------------------------- storage types
struct bf16s {
  uint16_t bits;
};

__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(bf16s x) { return __uint_as_float((uint32_t)x.bits << 16); }
__device__ __forceinline__ float to_f32(_Float16 x) { return (float)x; }

template <typename TS> __device__ __forceinline__ TS zero_of();
template <> __device__ __forceinline__ float zero_of<float>() { return 0.0f; }
template <> __device__ __forceinline__ bf16s zero_of<bf16s>() { return bf16s{0}; }
template <> __device__ __forceinline__ _Float16 zero_of<_Float16>() { return (_Float16)0.0f; }

inline size_t dtype_bytes(int dt) { return dt == PYRONN_F32 ? 4 : 2; }
inline bool dtype_ok(int dt) { return dt == PYRONN_F32 || dt == PYRONN_BF16 || dt == PYRONN_F16; }

// ----------------------------------------------------------------- arithmetic
// Linear blend a + t (b - a) written as the reference writes it:
// __fmaf_rn(b, t, __fmaf_rn(a, -t, a)) (cone_projector_3D_CudaKernel.cu.cc:74-82).
__device__ __forceinline__ float lerp_ref(float a, float b, float t) {
  return fmaf(b, t, fmaf(a, -t, a));
}

// Clamp into the zero border [-1, n] with one v_med3_f32.  For every input that is not
// a signaling NaN this equals fminf(fmaxf(c, -1), n) (a quiet NaN gives -1, as the
// oracle's std::fmin(std::fmax()) does); arithmetic never produces a signaling NaN.
// The test suite checks the equality on the device for all 2^32 inputs.
__device__ __forceinline__ float clamp_apron(float c, float n) {
  return __builtin_amdgcn_fmed3f(c, -1.0f, n);
}

// 1.0f / w, correctly rounded, in about half the instructions of the generic IEEE
// division.  For |w| in [2^-125, 2^125] (reciprocal and residuals stay normal): the
// hardware estimate v_rcp_f32 and two Newton-Raphson corrections
// r' = r + r (1 - w r), each residual formed exactly by one FMA (Markstein's
// correction step).  Everything else (zero, infinities, NaN, extreme exponents) takes
// the IEEE division.  Bit equality with 1.0f / w over all 2^32 inputs is checked on
// the device by the test suite.
__device__ __forceinline__ bool rcp_fast_ok(float w) {
  const float a = fabsf(w);
  return a >= 0x1p-125f && a <= 0x1p125f;
}
// Requires rcp_fast_ok(w).
__device__ __forceinline__ float rcp_fast(float w) {
  float r = __builtin_amdgcn_rcpf(w);
  r = fmaf(fmaf(-w, r, 1.0f), r, r);
  return fmaf(fmaf(-w, r, 1.0f), r, r);
}
__device__ __forceinline__ float rcp_rn(float w) { return rcp_fast_ok(w) ? rcp_fast(w) : 1.0f / w; }

// Iteration cap of every ray march (the reference has none and would spin
// forever when the step is below the float resolution of alpha).  At most 2^25
// iterations: with a step of at least half a voxel and alpha below 2^23 voxels, no
// ray of any volume the host accepts (W + 2 < 2^23, (H + 2)(D + 3) < 2^23) reaches it.
__device__ __forceinline__ int march_cap(float amin, float amax, float step) {
  if (!(amax > amin)) return 0;
  const float n = ceilf(1.25f * (amax - amin) / step) + 16.0f;
  return n < 33554432.0f ? (int)n : 33554432;
}

// ----------------------------------------------------------------- host errors
void set_error(const char *fmt, ...);
void clear_error();
int hip_fail(hipError_t e, const char *what);

#define PYRONN_CHECK_ARG(cond, ...)                        \
  do {                                                    \
    if (!(cond)) {                                        \
      ::pyronn::set_error(__VA_ARGS__);                   \
      return PYRONN_ERR_INVALID_ARGUMENT;                 \
    }                                                     \
  } while (0)

#define PYRONN_CHECK_LAUNCH(what)                                           \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess) return ::pyronn::hip_fail(e_, what);             \
  } while (0)

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace pyronn
