// Test kit: exhaustive device checks of the library's arithmetic shortcuts:
//  * rcp_rn() against the IEEE division 1.0f / w;
//  * clamp_apron() (one v_med3_f32) against the clamp fminf(fmaxf(c, -1), n) that the
//    oracle applies, with the reference value formed from the bit patterns by integer
//    compares (so no compiler rewrite can turn it into the instruction under test).
// Built by pyro-nn-layers_amd/build.py into tests/native/libpyronn_testkit.so with the
// library's own compiler flags, so the comparison sees the same code generation.
#include "common.h"

namespace {

// Thread t checks the 256 bit patterns t * 256 + k; the grid of 2^16 blocks of 256
// threads covers all 2^32 patterns.  counts[t] = number of patterns whose results
// differ (NaN results compare equal to NaN results).
__global__ void __launch_bounds__(256) rcp_check_kernel(unsigned *__restrict__ counts) {
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  unsigned bad = 0;
  for (unsigned k = 0; k < 256u; ++k) {
    const float w = __uint_as_float(t * 256u + k);
    const float a = pyronn::rcp_rn(w), b = 1.0f / w;
    const bool same = (__float_as_uint(a) == __float_as_uint(b)) || (a != a && b != b);
    bad += same ? 0u : 1u;
  }
  counts[t] = bad;
}

// Order-preserving integer key of a non-NaN float (-0 sorts just below +0).
__device__ __forceinline__ unsigned fkey(unsigned b) { return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }

// counts[t] = patterns t * 256 + k whose clamp differs from the reference; signaling
// NaN inputs (exponent all ones, quiet bit clear, payload non-zero) are excluded, as no
// arithmetic produces them.
__global__ void __launch_bounds__(256) clamp_check_kernel(unsigned *__restrict__ counts, float n) {
  const unsigned t = blockIdx.x * 256u + threadIdx.x;
  const unsigned klo = fkey(__float_as_uint(-1.0f)), khi = fkey(__float_as_uint(n));
  unsigned bad = 0;
  for (unsigned k = 0; k < 256u; ++k) {
    const unsigned b = t * 256u + k;
    const bool nan = (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu) != 0u;
    if (nan && (b & 0x00400000u) == 0u) continue;  // signaling NaN
    unsigned ref;
    if (nan) {
      ref = __float_as_uint(-1.0f);                  // fmax(qNaN, -1) = -1
    } else {
      const unsigned kc = fkey(b);
      ref = kc < klo ? __float_as_uint(-1.0f) : (kc > khi ? __float_as_uint(n) : b);
    }
    bad += __float_as_uint(pyronn::clamp_apron(__uint_as_float(b), n)) == ref ? 0u : 1u;
  }
  counts[t] = bad;
}

}  // namespace

extern "C" int pyronn_testkit_clamp_check(unsigned *counts, float n, void *stream) {
  hipLaunchKernelGGL(clamp_check_kernel, dim3(65536), dim3(256), 0, (hipStream_t)stream, counts, n);
  return (int)hipGetLastError();
}

// counts: device buffer of 2^24 unsigned ints.
extern "C" int pyronn_testkit_rcp_check(unsigned *counts, void *stream) {
  hipLaunchKernelGGL(rcp_check_kernel, dim3(65536), dim3(256), 0, (hipStream_t)stream, counts);
  return (int)hipGetLastError();
}
