// Per-CU issue rates of the two units that bound the cone kernels, measured on the device:
//  * vector-memory loads (the texture-address path): cycles per wave-wide load instruction
//    for address patterns that touch different numbers of 128-byte lines per group of four
//    lanes, from a buffer that stays in the L1 cache;
//  * VALU: lane operations per clock per CU for FP32 FMAs (three- and two-operand forms),
//    the lerp of the samplers, multiplies and floors, with 1-8 waves per SIMD.
// Prints one table line per case.  Built by pyro-nn-layers_amd/build.py next to the test
// kit; the rates explain the limiter figures the bench reports.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

namespace {

// Address patterns (offsets in 16-byte units; every access is masked into the buffer):
enum Pattern { kSame, kContiguous, kLinePerLane, kQuadPerLine, kPairPerLine, kActiveLanes, kHalfLinePerLane };

template <int BYTES>
__global__ void __launch_bounds__(256) gather_kernel(const uint4 *__restrict__ buf, int pattern, int active,
                                                     int iters, unsigned mask, uint4 *__restrict__ out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned off;
  switch (pattern) {
    case kSame: off = 0; break;
    case kContiguous: off = lane; break;                       // 64 x 16 B: 8 lines, one per two quads
    case kLinePerLane: off = lane * 8; break;                  // 64 lines, 4 per quad
    case kQuadPerLine: off = (lane >> 2) * 8 + (lane & 3); break;  // 16 lines, 1 per quad
    case kPairPerLine: off = (lane >> 1) * 8; break;           // 32 lines, 2 per quad
    case kActiveLanes: off = lane * 8; break;                  // lanes < active only, own line each
    default: off = lane * 4; break;                            // 64-byte stride: 2 lines per quad
  }
  off += (blockIdx.x * 4 + wave) * 512;
  uint4 acc = make_uint4(0, 0, 0, 0);
  const bool on = pattern != kActiveLanes || lane < active;
  for (int it = 0; it < iters; ++it) {
    const unsigned o = (off + (unsigned)it * 64u) & mask;
    if (on) {
      if (BYTES == 16) {
        const uint4 v = buf[o];
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
      } else {
        const uint2 v = reinterpret_cast<const uint2 *>(buf)[o * 2];
        acc.x ^= v.x;
        acc.y ^= v.y;
      }
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc;
}

enum ValuKind { kFma3, kFmac, kLerp, kMulFloor };

template <int KIND>
__global__ void __launch_bounds__(256) valu_kernel(float *__restrict__ out, int iters, float a, float b) {
  // loop-invariant operands in VGPRs (derived from the lane id so they are not scalars)
  const float av = a + 1e-30f * threadIdx.x, bv = b + 1e-30f * threadIdx.x;
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = (float)(threadIdx.x + k);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (KIND == kFma3) x[k] = fmaf(x[k], av, bv);          // v_fma_f32, three VGPR operands
      else if (KIND == kFmac) x[k] = fmaf(av, bv, x[k]);     // accumulator form (v_fmac_f32)
      else if (KIND == kLerp) x[k] = fmaf(bv, av, fmaf(x[k], -av, x[k]));  // the samplers' lerp
      else x[k] = floorf(x[k] * av);                         // v_mul_f32 + v_floor_f32
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int main() {
  int clk_khz = 0, cus = 0;
  CHECK(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const double clk = clk_khz * 1e3;
  std::printf("# %d CUs, nominal clock %.0f MHz (cycles below are at this clock)\n", cus, clk / 1e6);
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));

  // ---- loads: 16 KiB buffer (L1-resident), 8 blocks of 4 waves per CU
  const unsigned nelem = 1u << 10, mask = nelem - 1;
  const int gblocks = cus * 8, giters = 4096;
  uint4 *buf = nullptr, *gout = nullptr;
  CHECK(hipMalloc(&buf, nelem * sizeof(uint4)));
  CHECK(hipMemset(buf, 1, nelem * sizeof(uint4)));
  CHECK(hipMalloc(&gout, (size_t)gblocks * 256 * sizeof(uint4)));
  struct Case { int pattern, active; const char *name; int lines_per_quad; };
  const Case cases[] = {
      {kSame, 64, "all lanes one address", 1},           {kContiguous, 64, "contiguous 16-B pieces", 1},
      {kQuadPerLine, 64, "one line per quad", 1},         {kHalfLinePerLane, 64, "64-B stride", 2},
      {kPairPerLine, 64, "two lanes per line", 2},        {kLinePerLane, 64, "one line per lane", 4},
      {kActiveLanes, 32, "32 lanes, one line each", 4},   {kActiveLanes, 16, "16 lanes, one line each", 4},
      {kActiveLanes, 4, "4 lanes, one line each", 4},     {kActiveLanes, 1, "1 lane", 1},
  };
  std::printf("# load width, pattern, lines per active quad: cycles per wave load instruction per CU\n");
  for (int bytes : {16, 8}) {
    for (const Case &c : cases) {
      float ms = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {
        CHECK(hipEventRecord(e0));
        if (bytes == 16)
          gather_kernel<16><<<gblocks, 256>>>(buf, c.pattern, c.active, giters, mask, gout);
        else
          gather_kernel<8><<<gblocks, 256>>>(buf, c.pattern, c.active, giters, mask, gout);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&ms, e0, e1));
      }
      const double instr_per_cu = (double)gblocks * 4 * giters / cus;
      std::printf("load %2d B  %-26s %d lines/quad: %6.2f cycles\n", bytes, c.name, c.lines_per_quad,
                  ms * 1e-3 * clk / instr_per_cu);
    }
  }
  CHECK(hipFree(buf));
  CHECK(hipFree(gout));

  // ---- VALU
  const int viters = 1 << 14;
  float *vout = nullptr;
  CHECK(hipMalloc(&vout, (size_t)cus * 8 * 256 * sizeof(float)));
  const char *vnames[] = {"fma (3 VGPR operands)", "fma accumulator (fmac)", "lerp fma(b,t,fma(a,-t,a))",
                          "mul + floor"};
  const int ops_per_iter[] = {8, 8, 16, 16};
  std::printf("# VALU kind, waves per SIMD: lane operations per clock per CU\n");
  for (int kind = 0; kind < 4; ++kind) {
    for (int wps : {1, 2, 4, 8}) {
      const int blocks = cus * wps;  // 4 waves per block: wps waves per SIMD
      float ms = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {
        CHECK(hipEventRecord(e0));
        if (kind == kFma3) valu_kernel<kFma3><<<blocks, 256>>>(vout, viters, 0.999f, 0.001f);
        else if (kind == kFmac) valu_kernel<kFmac><<<blocks, 256>>>(vout, viters, 0.999f, 0.001f);
        else if (kind == kLerp) valu_kernel<kLerp><<<blocks, 256>>>(vout, viters, 0.999f, 0.001f);
        else valu_kernel<kMulFloor><<<blocks, 256>>>(vout, viters, 0.999f, 0.001f);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&ms, e0, e1));
      }
      const double lane_ops = (double)blocks * 256 * viters * ops_per_iter[kind];
      std::printf("valu %-28s %d waves/SIMD: %6.1f lane-ops per clock per CU\n", vnames[kind], wps,
                  lane_ops / cus / (ms * 1e-3 * clk));
    }
  }
  CHECK(hipFree(vout));
  return 0;
}
