// Cone-beam projector / backprojector for gfx950 (CDNA4).
//
// Kernels
//   cone_setup_kernel   per-view FP64 geometry (M^-1, camera centre) on the device
//   build_cells_kernel  copy of the volume as neighbourhood cells with a zero border
//                       (what the forward sampler reads)
//   build_det_cells_kernel  copy of the sinogram as 2x2 neighbourhood cells with a
//                       zero border (what the backprojection sampler reads)
//   cone_fwd_kernel     ray-driven forward projection, one lane per detector pixel
//   cone_bp_kernel      voxel-driven backprojection, one lane per ZPT voxels in z
//
// The forward march and the backprojection sum follow the reference kernels
// operation by operation (see the citations at each kernel), so that results
// agree with the host restatement to the last bit.
#include "common.h"
#include <mutex>
#include <type_traits>

namespace pyronn {

constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// =====================================================================================
// Per-view geometry (replaces cone_projector_3D_OPKernel.cc:96-182 and
// helper_geometry_cpu.h:30-43).  FP64 closed form: C = -M^-1 p4 is the right null
// vector of P with C(3) = 1; M^-1 = adj(M) / det(M).
// =====================================================================================
__global__ void cone_setup_kernel(const float *__restrict__ P, const float *__restrict__ origin,
                                  const float *__restrict__ spacing, int batch, int n,
                                  float *__restrict__ inv_ar, float *__restrict__ src) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= batch * n) return;
  const int b = idx / n;
  double m[12];
  for (int i = 0; i < 12; ++i) m[i] = (double)P[(size_t)idx * 12 + i];
  const double a = m[0], bb = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9],
               k = m[10];
  const double A = e * k - f * h, B = f * g - d * k, C = d * h - e * g;
  const double det = a * A + bb * B + c * C;
  double inv[9], s[3];
  bool ok = (fabs(det) > 1e-300) && isfinite(det);
  if (ok) {
    const double adj[9] = {A, c * h - bb * k, bb * f - c * e, B, a * k - c * g,
                           c * d - a * f, C, bb * g - a * h, a * e - bb * d};
    double mi[9];
    for (int i = 0; i < 9; ++i) mi[i] = adj[i] / det;
    const double p4[3] = {m[3], m[7], m[11]};
    const double sp[3] = {(double)spacing[b * 3 + 2], (double)spacing[b * 3 + 1],
                          (double)spacing[b * 3 + 0]};
    const double og[3] = {(double)origin[b * 3 + 2], (double)origin[b * 3 + 1],
                          (double)origin[b * 3 + 0]};
    for (int r = 0; r < 3; ++r) {
      const double cr = -(mi[r * 3 + 0] * p4[0] + mi[r * 3 + 1] * p4[1] + mi[r * 3 + 2] * p4[2]);
      s[r] = (cr - og[r]) / sp[r];
      for (int cc = 0; cc < 3; ++cc) inv[r * 3 + cc] = mi[r * 3 + cc] / sp[r];
    }
    for (int i = 0; i < 9; ++i) ok = ok && isfinite(inv[i]);
    for (int i = 0; i < 3; ++i) ok = ok && isfinite(s[i]);
  }
  const float qnan = __builtin_nanf("");
  for (int i = 0; i < 9; ++i) inv_ar[(size_t)idx * 9 + i] = ok ? (float)inv[i] : qnan;
  for (int i = 0; i < 3; ++i) src[(size_t)idx * 3 + i] = ok ? (float)s[i] : qnan;
}

// =====================================================================================
// Neighbourhood cells.  The forward projector reads the volume through a copy made
// of cells that hold a voxel neighbourhood, with zeros outside the volume, stored
// z-fastest so that the 64 rays of a wave (16 detector rows ~ 16 neighbouring z
// positions, and z is transverse to every ray of an orbit about the z axis) fall into few cache lines
// for every view direction.  The zero border makes the reference's per-neighbour
// bounds tests unnecessary.  Layout (xy quads): cell (x, y, z), x in [-1, W],
// y in [-1, H], z in [-1, D + 1], holds [v(x,y,z), v(x+1,y,z), v(x,y+1,z), v(x+1,y+1,z)]
// at index ((x+1) (H+2) + (y+1)) (D+3) + (z+1).  A sample needs the quads at z and
// z + 1, which are adjacent:
//  * f32 - 16-byte quads; a sample is two adjacent 16-byte loads sharing cache lines;
//  * bf16 / f16 - 8-byte quads; a sample is one 16-byte load at 8-byte alignment.
// Every storage type thus spends one quad per z step of a column: the 16-bit copy is
// half the bytes of the f32 one.  (A 2x2x2 cell per z step for 16-bit storage, also one
// 16-byte load per sample, occupies as many bytes as the f32 quads: on the 512^3 /
// 720-view forward it took 473 ms against 389 ms for these quads in bf16, as the 16-bit
// gathers are limited by L2 misses rather than by instruction issue.)
// The build stages (x, z) tiles of two neighbouring y rows of the volume in LDS, so that
// it reads the volume along x and writes the cells along z, both contiguously, with no
// intermediate copy of the volume.
// =====================================================================================
template <typename TS> struct CellLayout {
  static constexpr int kPerCell = 4;  // elements per cell (an xy quad)
  static constexpr int kZExtra = 3;   // cells along z beyond D
};

// The 16 bytes of the z and z + 1 quads of a 16-bit cell array: 8-byte aligned (the
// target supports unaligned vector loads, so this stays one global_load_dwordx4).
typedef uint32_t u32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));

template <typename TS> static size_t cell_elems(int D, int H, int W) {
  return (size_t)(W + 2) * (H + 2) * (D + CellLayout<TS>::kZExtra) * CellLayout<TS>::kPerCell;
}

// The four neighbours of a cell in one store: 16 bytes (f32) or 8 bytes (16-bit types).
__device__ __forceinline__ void store_cell(float *o, float a, float b, float c, float d) {
  *reinterpret_cast<float4 *>(o) = make_float4(a, b, c, d);
}
template <typename TS>
__device__ __forceinline__ void store_cell(TS *o, TS a, TS b, TS c, TS d) {
  static_assert(sizeof(TS) == 2, "16-bit storage");
  uint16_t h[4];
  __builtin_memcpy(&h[0], &a, 2);
  __builtin_memcpy(&h[1], &b, 2);
  __builtin_memcpy(&h[2], &c, 2);
  __builtin_memcpy(&h[3], &d, 2);
  *reinterpret_cast<uint2 *>(o) =
      make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
}

// One tile = cells (cx0 .. cx0 + 63, cy, cz0 .. cz0 + 31) in cell coordinates (voxel
// x = cx - 1, y = cy - 1, z = cz - 1).  The block reads the voxels x = cx0 - 1 .. cx0 + 63
// of rows y and y + 1 at the tile's 32 z positions into LDS (one wave row per (y, z),
// contiguous in x; zero outside the volume), then writes each cell's quad, lanes along
// z (32 consecutive 16- or 8-byte cells per half wave).  Tiles are numbered with cy
// fastest, so blocks running together read overlapping rows.
constexpr int kCellTileX = 64, kCellTileZ = 32;

template <typename TS>
__global__ void __launch_bounds__(256)
build_cells_kernel(const TS *__restrict__ vol, TS *__restrict__ cells, int W, int H, int D,
                   int batch) {
  using L = CellLayout<TS>;
  __shared__ TS t[2][kCellTileZ][kCellTileX + 1];
  const int Wc = W + 2, Hc = H + 2, Dc = D + L::kZExtra;
  const long long ntx = (Wc + kCellTileX - 1) / kCellTileX, ntz = (Dc + kCellTileZ - 1) / kCellTileZ;
  const long long tiles = ntx * ntz * Hc * batch;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t vstride = (size_t)W * H * D;
  for (long long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    long long r = tile;
    const int cy = (int)(r % Hc); r /= Hc;
    const int tx = (int)(r % ntx); r /= ntx;
    const int tz = (int)(r % ntz);
    const long long b = r / ntz;
    const int cx0 = tx * kCellTileX, cz0 = tz * kCellTileZ;
    const TS *v = vol + (size_t)b * vstride;
    // 64 rows of (y, z): wave w loads rows w, w + 4, ...; lane l loads voxel x = cx0 - 1 + l,
    // lane 0 also the 65th voxel x = cx0 + 63
    for (int row = wave; row < 2 * kCellTileZ; row += 4) {
      const int ry = row / kCellTileZ, rz = row % kCellTileZ;
      const int y = cy - 1 + ry, z = cz0 - 1 + rz;
      const bool yz_in = y >= 0 && y < H && z >= 0 && z < D;
      const TS *vr = v + ((size_t)z * H + y) * W;
      const int x = cx0 - 1 + lane;
      t[ry][rz][lane] = (yz_in && x >= 0 && x < W) ? vr[x] : zero_of<TS>();
      if (lane == 0) {
        const int x2 = cx0 + kCellTileX - 1;
        t[ry][rz][kCellTileX] = (yz_in && x2 < W) ? vr[x2] : zero_of<TS>();
      }
    }
    __syncthreads();
    const int j = threadIdx.x & 31;
    const int cz = cz0 + j;
#pragma unroll
    for (int k = 0; k < kCellTileX / 8; ++k) {
      const int i = (threadIdx.x >> 5) + 8 * k;
      const int cx = cx0 + i;
      if (cx < Wc && cz < Dc) {
        TS *o = cells + ((((size_t)b * Wc + cx) * Hc + cy) * Dc + cz) * L::kPerCell;
        store_cell(o, t[0][j][i], t[0][j][i + 1], t[1][j][i], t[1][j][i + 1]);
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void unpack2(uint32_t w, bf16s, float &a, float &b) {
  a = __uint_as_float(w << 16);
  b = __uint_as_float(w & 0xffff0000u);
}
__device__ __forceinline__ void unpack2(uint32_t w, _Float16, float &a, float &b) {
  _Float16 h[2];
  __builtin_memcpy(h, &w, 4);
  a = (float)h[0];
  b = (float)h[1];
}

// The eight corners of the cell a lane read last.  Consecutive samples of a ray are
// half a voxel apart, so about half of them fall in the cell of the sample before;
// those reuse the registers instead of loading again.  The blend is the same
// arithmetic on the same values either way.  The all-ones key is never a cell index
// (the 32-bit index serves cell counts below 2^31, the 64-bit one counts below 2^63),
// so the first sample of a ray always loads.
// f32 and bf16 keep the corners as floats (bf16 widened once per load).  f16 keeps the
// packed cell: the blend's FMAs read the halves directly (v_fma_mix_f32 widens an f16
// operand exactly inside the FMA, one rounding as before), so a load costs no
// conversions at all.
// IT is the index type of the cell array: int for arrays within the 32-bit limits of
// fwd_small_index, long long beyond them.  The 16-bit streamed march also takes
// IT = unsigned: 32-bit byte offsets from a biased base (locate16, fwd_byte_offsets).
template <typename IT> struct CellKey { typedef unsigned type; };
template <> struct CellKey<long long> { typedef unsigned long long type; };

template <typename TS, typename IT> struct CellRegs {
  typename CellKey<IT>::type cell = ~(typename CellKey<IT>::type)0;
  float c[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  __device__ __forceinline__ float corner(int i) const { return c[i]; }
};
template <typename IT> struct CellRegs<_Float16, IT> {
  typename CellKey<IT>::type cell = ~(typename CellKey<IT>::type)0;
  uint32_t w[4] = {0u, 0u, 0u, 0u};  // corners 2k (low half) and 2k + 1 (high half)
  __device__ __forceinline__ float corner(int i) const {
    _Float16 h[2];
    __builtin_memcpy(h, &w[i >> 1], 4);
    return (float)h[i & 1];
  }
};

// Trilinear sample at (px, py, pz), lerp z -> y -> x exactly as
// cone_projector_3D_CudaKernel.cu.cc:74-82.  With CLAMP the coordinate is clamped into
// the zero border first (identical results for every finite coordinate); without it the
// caller guarantees every coordinate lies in [-1, N + 1).
// Dc = cells along z, HDc = (H + 2) * Dc: strides of the z-fastest cell array.
//  * IT = int: the (y, z) part of the index, fy * Dc + fz, is an integer of magnitude
//    below 2^23 (the host keeps HDc below 2^23), so one float FMA forms it exactly; the
//    x part uses a signed 24-bit multiply (the host keeps W + 2 below 2^23 and the cell
//    count below 2^31).
//  * IT = long long: 64-bit integer products, for any volume whose sides are below 2^24
//    (so that every coordinate and its floor are exact floats).
template <typename TS, bool CLAMP, typename IT>
__device__ __forceinline__ float sample_cells(const TS *__restrict__ cells, IT Dc, IT HDc,
                                              float px, float py, float pz, float fW, float fH,
                                              float fD, CellRegs<TS, IT> &r) {
  if (CLAMP) {
    px = clamp_apron(px, fW);
    py = clamp_apron(py, fH);
    pz = clamp_apron(pz, fD);
  }
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  const float xd = px - fx, yd = py - fy, zd = pz - fz;
  typename CellKey<IT>::type cell;
  if constexpr (sizeof(IT) == 4) {
    const int yz = (int)fmaf(fy, (float)Dc, fz);
    cell = (unsigned)(__mul24((int)fx, HDc) + yz + (HDc + Dc + 1));
  } else {
    cell = (unsigned long long)((long long)(int)fx * HDc + (long long)(int)fy * Dc + (long long)(int)fz +
                                (HDc + Dc + 1));
  }
  if (cell != r.cell) {
    r.cell = cell;
    if constexpr (sizeof(TS) == 4) {
      const float4 *c4 = reinterpret_cast<const float4 *>(cells) + cell;
      const float4 a = c4[0], b = c4[1];  // quads at z and z + 1
      r.c[0] = a.x; r.c[1] = a.y; r.c[2] = a.z; r.c[3] = a.w;
      r.c[4] = b.x; r.c[5] = b.y; r.c[6] = b.z; r.c[7] = b.w;
    } else {
      // quads at z and z + 1: words 0, 1 and 2, 3
      const u32x4_a8 q = *reinterpret_cast<const u32x4_a8 *>(reinterpret_cast<const uint2 *>(cells) + cell);
      if constexpr (std::is_same<TS, _Float16>::value) {
        r.w[0] = q.x; r.w[1] = q.y; r.w[2] = q.z; r.w[3] = q.w;
      } else {
        unpack2(q.x, TS{}, r.c[0], r.c[1]);
        unpack2(q.y, TS{}, r.c[2], r.c[3]);
        unpack2(q.z, TS{}, r.c[4], r.c[5]);
        unpack2(q.w, TS{}, r.c[6], r.c[7]);
      }
    }
  }
  // corners: p000 p001 p010 p011 (z) then the same four at z + 1; digits are z, y, x.
  const float p00 = lerp_ref(r.corner(0), r.corner(4), zd);
  const float p01 = lerp_ref(r.corner(1), r.corner(5), zd);
  const float p10 = lerp_ref(r.corner(2), r.corner(6), zd);
  const float p11 = lerp_ref(r.corner(3), r.corner(7), zd);
  const float p0 = lerp_ref(p00, p10, yd);
  const float p1 = lerp_ref(p01, p11, yd);
  return lerp_ref(p0, p1, xd);
}

// Trilinear sample read straight from the volume [D][H][W], with no cell copy (the
// forward variant that allocates nothing).  The eight neighbours of the cell that
// sample_cells would read are gathered with a bounds test each (zero outside the volume,
// as the cell array's border holds) and blended in the same order, so the two samplers
// give the same bits for every coordinate.  W and HW are the row and plane strides.
template <typename TS, bool CLAMP>
__device__ __forceinline__ float sample_direct(const TS *__restrict__ vol, long long W, long long HW,
                                               float px, float py, float pz, float fW, float fH,
                                               float fD) {
  if (CLAMP) {
    px = clamp_apron(px, fW);
    py = clamp_apron(py, fH);
    pz = clamp_apron(pz, fD);
  }
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  const float xd = px - fx, yd = py - fy, zd = pz - fz;
  // the floors lie in [-1, N] (clamped, or proven by the caller): exact small integers
  const bool x0 = fx >= 0.0f && fx < fW, x1 = fx >= -1.0f && fx + 1.0f < fW;
  const bool y0 = fy >= 0.0f && fy < fH, y1 = fy >= -1.0f && fy + 1.0f < fH;
  const bool z0 = fz >= 0.0f && fz < fD, z1 = fz >= -1.0f && fz + 1.0f < fD;
  const long long base = (long long)(int)fz * HW + (long long)(int)fy * W + (long long)(int)fx;
  const TS *p = vol + base;
  const float zero = 0.0f;
  // corner digits are z, y, x, as in sample_cells
  const float c000 = (z0 && y0 && x0) ? to_f32(p[0]) : zero;
  const float c001 = (z0 && y0 && x1) ? to_f32(p[1]) : zero;
  const float c010 = (z0 && y1 && x0) ? to_f32(p[W]) : zero;
  const float c011 = (z0 && y1 && x1) ? to_f32(p[W + 1]) : zero;
  const float c100 = (z1 && y0 && x0) ? to_f32(p[HW]) : zero;
  const float c101 = (z1 && y0 && x1) ? to_f32(p[HW + 1]) : zero;
  const float c110 = (z1 && y1 && x0) ? to_f32(p[HW + W]) : zero;
  const float c111 = (z1 && y1 && x1) ? to_f32(p[HW + W + 1]) : zero;
  const float p00 = lerp_ref(c000, c100, zd);
  const float p01 = lerp_ref(c001, c101, zd);
  const float p10 = lerp_ref(c010, c110, zd);
  const float p11 = lerp_ref(c011, c111, zd);
  const float p0 = lerp_ref(p00, p10, yd);
  const float p1 = lerp_ref(p01, p11, yd);
  return lerp_ref(p0, p1, xd);
}

// DIRECT: c is the volume itself and Dc, HDc its row and plane strides (sample_direct);
// otherwise c is the cell array (sample_cells).
template <typename TS, bool HW, bool CLAMP, typename IT, bool DIRECT>
__device__ __forceinline__ float cone_sample(const TS *__restrict__ c, IT Dc, IT HDc, float px,
                                             float py, float pz, float fW, float fH, float fD,
                                             CellRegs<TS, IT> &r) {
  if (HW) {  // tex3D(p + 0.5): the sampler subtracts the half texel again
    px = (px + 0.5f) - 0.5f;
    py = (py + 0.5f) - 0.5f;
    pz = (pz + 0.5f) - 0.5f;
  }
  if constexpr (DIRECT)
    return sample_direct<TS, CLAMP>(c, (long long)Dc, (long long)HDc, px, py, pz, fW, fH, fD);
  else
    return sample_cells<TS, CLAMP, IT>(c, Dc, HDc, px, py, pz, fW, fH, fD, r);
}

// The 16-bit forward's march, streamed: every sample loads its cell, and the load of the
// next sample is issued before the current sample is blended, so the blend's arithmetic
// covers the load's latency.  (The 16-bit forward issues one load per march iteration and
// is bound by latency and VALU issue rather than by the address path; the f32 forward,
// which issues two and is bound by the address path, keeps march_ray.)  Same samples, same
// positions, same blend on the same values, accumulated in the same order as march_ray:
// the tail's re-evaluation of the last sample is that sample's value, reused.
template <typename TS, bool HW, bool CLAMP, typename IT>
__device__ __forceinline__ void locate16(float px, float py, float pz, IT Dc, IT HDc, float fW, float fH, float fD,
                                         typename CellKey<IT>::type &cell, float &xd, float &yd, float &zd) {
  if (HW) {  // tex3D(p + 0.5): the sampler subtracts the half texel again
    px = (px + 0.5f) - 0.5f;
    py = (py + 0.5f) - 0.5f;
    pz = (pz + 0.5f) - 0.5f;
  }
  if (CLAMP) {
    px = clamp_apron(px, fW);
    py = clamp_apron(py, fH);
    pz = clamp_apron(pz, fD);
  }
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  xd = px - fx;
  yd = py - fy;
  zd = pz - fz;
  if constexpr (std::is_same<IT, unsigned>::value) {
    // Byte offset of the cell, two VALU instructions after the floors.  yz = fy * Dc + fz
    // is an exact integer-valued float in [-(Dc + 1), 2^20); m = 8 yz + 2^23 + 8 (Dc + 1)
    // is an exact integer in [2^23, 2^24) (the host keeps 8 HDc below 2^23), so its bit
    // pattern is 0x4B000000 + m - 2^23, and the 24-bit multiply-add puts the x part on
    // top.  The result, byte offset + 0x4B000000 - 8 HDc, lies in [0, 2^32) (the host
    // keeps the cell array below 0xB5000000 bytes); cell_base subtracts the bias.
    const float yz = fmaf(fy, (float)Dc, fz);
    const float m = fmaf(yz, 8.0f, 0x1p23f + (float)(8u * (Dc + 1u)));
    cell = (unsigned)__mul24((int)fx, (int)(8u * HDc)) + __float_as_uint(m);
  } else if constexpr (sizeof(IT) == 4) {
    const int yz = (int)fmaf(fy, (float)Dc, fz);
    cell = (unsigned)(__mul24((int)fx, HDc) + yz + (HDc + Dc + 1));
  } else {
    cell = (unsigned long long)((long long)(int)fx * HDc + (long long)(int)fy * Dc + (long long)(int)fz +
                                (HDc + Dc + 1));
  }
}

// The address locate16's cell keys count from: the cell array itself, or for IT = unsigned
// (byte offsets) the array moved by 8 HDc - 0x4B000000 bytes (wrapping 64-bit arithmetic;
// every load adds the offset back).
template <typename TS, typename IT>
__device__ __forceinline__ const char *cell_base(const TS *cells, IT HDc) {
  if constexpr (std::is_same<IT, unsigned>::value)
    return reinterpret_cast<const char *>(cells) + (8ll * HDc - 0x4B000000ll);
  else
    return reinterpret_cast<const char *>(cells);
}

template <typename IT, typename K>
__device__ __forceinline__ u32x4_a8 load16(const char *__restrict__ base, K cell) {
  if constexpr (std::is_same<IT, unsigned>::value)
    return *reinterpret_cast<const u32x4_a8 *>(base + cell);
  else
    return *reinterpret_cast<const u32x4_a8 *>(reinterpret_cast<const uint2 *>(base) + cell);
}

// The trilinear blend of sample_cells on a loaded 16-bit cell (quads at z and z + 1).
template <typename TS>
__device__ __forceinline__ float blend16(const u32x4_a8 q, float xd, float yd, float zd) {
  float c[8];
  if constexpr (std::is_same<TS, _Float16>::value) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      _Float16 h[2];
      __builtin_memcpy(h, &w[i >> 1], 4);
      c[i] = (float)h[i & 1];
    }
  } else {
    unpack2(q.x, TS{}, c[0], c[1]);
    unpack2(q.y, TS{}, c[2], c[3]);
    unpack2(q.z, TS{}, c[4], c[5]);
    unpack2(q.w, TS{}, c[6], c[7]);
  }
  const float p00 = lerp_ref(c[0], c[4], zd);
  const float p01 = lerp_ref(c[1], c[5], zd);
  const float p10 = lerp_ref(c[2], c[6], zd);
  const float p11 = lerp_ref(c[3], c[7], zd);
  const float p0 = lerp_ref(p00, p10, yd);
  const float p1 = lerp_ref(p01, p11, yd);
  return lerp_ref(p0, p1, xd);
}

template <typename TS, bool HW, bool GUARDED, typename IT>
__device__ __forceinline__ float march_ray_stream(const TS *__restrict__ v, IT Dc, IT HDc, float fW, float fH,
                                                  float fD, float rx, float ry, float rz, float s0, float s1,
                                                  float s2, float amin, float amax, float step) {
  const char *vb = cell_base<TS, IT>(v, HDc);
  typename CellKey<IT>::type cell;
  float xd, yd, zd;
  float pix = 0.0f, vlast = 0.0f;
  const int cap = GUARDED ? march_cap(amin, amax, step) : 0;
  if (amin < amax) {  // entrance sample, weight 1/2
    locate16<TS, HW, GUARDED, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH, fD,
                                  cell, xd, yd, zd);
    vlast = blend16<TS>(load16<IT>(vb, cell), xd, yd, zd);
    pix = fmaf(0.5f, vlast, pix);
    amin += step;
  }
  int it = 0;
  bool more = amin < amax && (!GUARDED || it < cap);
  // Two register sets, A and B, take turns: the loop body handles two samples, so that no
  // cell or offsets move between registers from one sample to the next.
  u32x4_a8 qa = {0u, 0u, 0u, 0u}, qb = {0u, 0u, 0u, 0u};
  float xa = 0.0f, ya = 0.0f, za = 0.0f, xb = 0.0f, yb = 0.0f, zb = 0.0f;
  if (more) {
    locate16<TS, HW, GUARDED, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH, fD,
                                  cell, xa, ya, za);
    qa = load16<IT>(vb, cell);
  }
  if constexpr (!GUARDED) {
    // Samples every lane of the wave is certain to take, marched first with no per-lane
    // test: a wave-uniform count, so the loop test is scalar and no lane is masked.  Each
    // alpha += step adds at most step + 2^-24 (|alpha| + step) while alpha stays within
    // [amin, amax] (one rounding of at most 2^-24 of the sum), so j increments stay below
    // amax while j (step + 2^-24 (M + step)) < amax - amin, M = max(|amin|, |amax|).  The
    // count below takes 2^-22 for 2^-24, which also covers the roundings of its own three
    // operations, and two increments less; the loop after it then runs the remaining
    // samples with the test, so the samples and their order are those of the tested loop.
    int safe = 0;
    if (more) {
      const float mag = fmaxf(fabsf(amin), fabsf(amax));
      const float adv = step + (mag + step) * 0x1p-22f;
      safe = (int)((amax - amin) / adv) - 2;
    }
    // (the reduction assumes every lane of the wave active, which fwd_block arranges; a
    // wave with inactive lanes would skip this phase)
    const int wmin = __ockl_wfred_min_i32(safe);
    const int nsafe = __builtin_amdgcn_read_exec() == ~0ull ? __builtin_amdgcn_readfirstlane(wmin) : 0;
    if (nsafe >= 4) {
      // Three register sets, A, B and C, take turns, so that two loads are in flight while
      // a sample is blended: B is requested first, then each step requests the sample two
      // ahead and blends the oldest.  1 + 3 t increments, t = (nsafe - 1) / 3.
      u32x4_a8 qc;
      float xc, yc, zc;
      amin += step;
      locate16<TS, HW, false, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH, fD,
                                  cell, xb, yb, zb);
      qb = load16<IT>(vb, cell);
      const int triples = (nsafe - 1) / 3;
      for (int k = 0; k < triples; ++k) {
        amin += step;
        locate16<TS, HW, false, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH,
                                    fD, cell, xc, yc, zc);
        qc = load16<IT>(vb, cell);
        pix += blend16<TS>(qa, xa, ya, za);
        amin += step;
        locate16<TS, HW, false, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH,
                                    fD, cell, xa, ya, za);
        qa = load16<IT>(vb, cell);
        pix += blend16<TS>(qb, xb, yb, zb);
        amin += step;
        locate16<TS, HW, false, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH,
                                    fD, cell, xb, yb, zb);
        qb = load16<IT>(vb, cell);
        pix += blend16<TS>(qc, xc, yc, zc);
      }
      // A and B hold the next two samples and alpha is B's: blend A, and B becomes A
      pix += blend16<TS>(qa, xa, ya, za);
      qa = qb;
      xa = xb;
      ya = yb;
      za = zb;
    }
  }
  while (more) {
    // the sample in A; the next sample's cell requested into B before A is blended
    amin += step;
    ++it;
    more = amin < amax && (!GUARDED || it < cap);
    if (more) {
      locate16<TS, HW, GUARDED, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH,
                                    fD, cell, xb, yb, zb);
      qb = load16<IT>(vb, cell);
    }
    vlast = blend16<TS>(qa, xa, ya, za);
    pix += vlast;
    if (!more) break;
    // the sample in B; the next one into A
    amin += step;
    ++it;
    more = amin < amax && (!GUARDED || it < cap);
    if (more) {
      locate16<TS, HW, GUARDED, IT>(fmaf(amin, rx, s0), fmaf(amin, ry, s1), fmaf(amin, rz, s2), Dc, HDc, fW, fH,
                                    fD, cell, xa, ya, za);
      qa = load16<IT>(vb, cell);
    }
    vlast = blend16<TS>(qb, xb, yb, zb);
    pix += vlast;
  }
  pix *= step;
  if (pix > 0.0f) {  // tail only for a positive sum (cone_projector_3D_CudaKernel.cu.cc:164-179)
    const float flast = vlast;
    pix = fmaf(-(0.5f * step), flast, pix);
    amin -= step;
    const float last = amax - amin;
    pix = fmaf(0.5f * last, flast, pix);
    locate16<TS, HW, GUARDED, IT>(fmaf(amax, rx, s0), fmaf(amax, ry, s1), fmaf(amax, rz, s2), Dc, HDc, fW, fH, fD,
                                  cell, xd, yd, zd);
    pix = fmaf(0.5f * last, blend16<TS>(load16<IT>(vb, cell), xd, yd, zd), pix);
  }
  return pix;
}

// The march of one ray from its clipped parameter interval: entrance sample with weight
// 1/2, loop samples, and the tail rule (cone_projector_3D_CudaKernel.cu.cc:134-179).
// GUARDED = false drops the coordinate clamp and the iteration cap; the caller uses it
// only for rays where neither can change anything (see fwd_block).
template <typename TS, bool HW, bool GUARDED, typename IT, bool DIRECT>
__device__ __forceinline__ float march_ray(const TS *__restrict__ v, IT Dc, IT HDc, float fW,
                                           float fH, float fD, float rx, float ry, float rz,
                                           float s0, float s1, float s2, float amin, float amax,
                                           float step) {
  float pix = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
  CellRegs<TS, IT> regs;
  const int cap = GUARDED ? march_cap(amin, amax, step) : 0;
  if (amin < amax) {
    px = fmaf(amin, rx, s0);
    py = fmaf(amin, ry, s1);
    pz = fmaf(amin, rz, s2);
    pix = fmaf(0.5f, cone_sample<TS, HW, GUARDED, IT, DIRECT>(v, Dc, HDc, px, py, pz, fW, fH, fD, regs), pix);
    amin += step;
  }
  for (int it = 0; amin < amax && (!GUARDED || it < cap); ++it) {
    px = fmaf(amin, rx, s0);
    py = fmaf(amin, ry, s1);
    pz = fmaf(amin, rz, s2);
    pix += cone_sample<TS, HW, GUARDED, IT, DIRECT>(v, Dc, HDc, px, py, pz, fW, fH, fD, regs);
    amin += step;
  }
  pix *= step;
  if (pix > 0.0f) {  // tail only for a positive sum (cone_projector_3D_CudaKernel.cu.cc:164-179)
    const float flast = cone_sample<TS, HW, GUARDED, IT, DIRECT>(v, Dc, HDc, px, py, pz, fW, fH, fD, regs);
    pix = fmaf(-(0.5f * step), flast, pix);
    amin -= step;
    const float last = amax - amin;
    pix = fmaf(0.5f * last, flast, pix);
    px = fmaf(amax, rx, s0);
    py = fmaf(amax, ry, s1);
    pz = fmaf(amax, rz, s2);
    pix = fmaf(0.5f * last, cone_sample<TS, HW, GUARDED, IT, DIRECT>(v, Dc, HDc, px, py, pz, fW, fH, fD, regs),
               pix);
  }
  return pix;
}

// =====================================================================================
// Forward projection.  One lane per (view, row, col); a 256-lane block covers a
// 16 x 16 detector tile and each wave 4 columns x 16 rows of it.  For an orbit about z,
// rays of one detector column stay at nearly the same (x, y) and differ in z, so with
// z-fastest cells the 16 rows of a wave column share cache lines (measured 3% faster
// than 8 x 8 waves at 512^3).  Lanes are numbered rows fastest: the texture-address
// unit takes a wave's load four lanes at a time and spends a cycle per distinct 128-byte
// line within each group of four, so each group is four neighbouring rows of one column
// (z-neighbours in one z-fastest cell column, mostly one line) rather than four columns
// (up to four different cell columns).
// Work order: workgroups are dispatched round-robin over the 8 XCDs, so block id
// b runs on XCD b % 8.  XCD k takes detector tile rows k, k+8, k+16, ...; within a
// tile row it walks tiles (outer) and views (inner).  The blocks resident on one XCD
// at any time are therefore one or two 16 x 16 tiles of one 16-row band over a few
// hundred consecutive views: their rays cross the same z slab of the volume, and the
// rays of one pixel in consecutive views are half a degree apart, so lines fetched
// for one view are read again by the next while they are still in that XCD's L2.
// (Measured on the 512^3 / 720-view forward against views outer, tiles inner: 483
// against 534 ms in f32, 328 against 390 ms in bf16; groups of 2 to 32 tiles per pass
// over the views fell in between, the fewer tiles the faster.)  The order changes
// only which lane computes which ray, never the arithmetic of a ray.
// Ray setup: cone_projector_3D_CudaKernel.cu.cc:197-221; march: :87-181
// (hardware_interp: cone_projector_3D_CudaKernel_hardware_interp.cu.cc:43-142).
// =====================================================================================
// One forward problem (one volume and its views) as the kernels see it.
struct FwdProblem {
  const void *cells;     // z-fastest neighbourhood cells of the volume
  float *sino;           // [n][H_d][W_d]
  const float *inv_ar;   // [n][9]
  const float *src;      // [n][3]
  const float *step;     // [1]
  const float *spacing;  // [3] (z, y, x)
  int W, H, D, n, Hd, Wd, tiles_x, tiles_y;
  long long begin;       // first work item of the problem in a batched launch (multiple of 8)
};

// Forward block shapes: a 256-lane block covers TC detector columns x TR rows and each of
// its four waves WC columns x WR = 64 / WC rows (waves side by side along the columns,
// then along the rows), lanes numbered rows fastest within a wave.  FwdTile16 is the
// shape of the single-problem launches; the batched launch takes any of kFwdTiles.
template <int TC_, int TR_, int WC_> struct FwdTile {
  static constexpr int TC = TC_, TR = TR_, WC = WC_, WR = 64 / WC_, WX = TC_ / WC_;
  static_assert(TC * TR == 256 && WC * WR == 64 && (TC / WC) * (TR / WR) == 4, "four 64-lane waves");
};
using FwdTile16 = FwdTile<16, 16, 4>;
constexpr int kFwdTiles = 4;  // 0: 16 x 16 (4 x 16 waves), 1: 8 x 32 (2 x 32), 2: 32 x 8 (8 x 8), 3: 16 x 16 (8 x 8)
__host__ __device__ constexpr int fwd_tile_cols(int t) { return t == 1 ? 8 : t == 2 ? 32 : 16; }
__host__ __device__ constexpr int fwd_tile_rows(int t) { return t == 1 ? 32 : t == 2 ? 8 : 16; }

// DIRECT: P.cells is the volume [D][H][W] itself, sampled with sample_direct.
template <typename TS, bool HW, typename IT, bool DIRECT = false, typename TL = FwdTile16>
__device__ __forceinline__ void fwd_block(const FwdProblem &P, long long q) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xcd = (int)(q % kXcds);
  const long long qq = q / kXcds;
  const long long per_row = (long long)P.n * P.tiles_x;
  const int band = (int)(qq / per_row);
  const int ty = band * kXcds + xcd;
  if (ty >= P.tiles_y) return;
  const long long w_ = qq - (long long)band * per_row;
  const int tx = (int)(w_ / P.n);
  const int view = (int)(w_ - (long long)tx * P.n);
  const int W = P.W, H = P.H, D = P.D, Hd = P.Hd, Wd = P.Wd;
  // Lanes past the detector's last column or row: in the 16-bit streamed march they march a
  // copy of the edge pixel's ray and store nothing, so that every lane of the wave takes
  // part in its wave-wide sample count; elsewhere they stop here.
  const int col0 = tx * TL::TC + (wave % TL::WX) * TL::WC + lane / TL::WR;
  const int row0 = ty * TL::TR + (wave / TL::WX) * TL::WR + lane % TL::WR;
  const bool store = col0 < Wd && row0 < Hd;
  if constexpr (sizeof(TS) == 4 || DIRECT)
    if (!store) return;
  const int col = col0 < Wd ? col0 : Wd - 1, row = row0 < Hd ? row0 : Hd - 1;

  const float *m = P.inv_ar + (size_t)view * 9;
  const float s0 = P.src[view * 3 + 0], s1 = P.src[view * 3 + 1], s2 = P.src[view * 3 + 2];
  const float step = P.step[0];
  const float spx = P.spacing[2], spy = P.spacing[1], spz = P.spacing[0];
  const TS *v = static_cast<const TS *>(P.cells);
  // z-fastest cell strides, or the volume's row and plane strides (DIRECT)
  const IT Wp = DIRECT ? (IT)W : (IT)D + CellLayout<TS>::kZExtra;
  const IT HWp = DIRECT ? (IT)H * W : (IT)(H + 2) * Wp;
  const float fW = (float)W, fH = (float)H, fD = (float)D;
  float *out = P.sino + ((size_t)view * Hd + row) * (size_t)Wd + col;
  const float u = (float)col, w = (float)row;
  float rx = fmaf(u, m[0], fmaf(w, m[1], m[2]));
  float ry = fmaf(u, m[3], fmaf(w, m[4], m[5]));
  float rz = fmaf(u, m[6], fmaf(w, m[7], m[8]));
  const float nrm = 1.0f / sqrtf(fmaf(rz, rz, fmaf(ry, ry, rx * rx)));
  rx *= nrm;
  ry *= nrm;
  rz *= nrm;

  if (!(step > 0.0f)) {
    if (store) *out = __builtin_nanf("");
    return;
  }

  // Slab clipping (cone_projector_3D_CudaKernel.cu.cc:96-132; texture box :53-79).
  const float lo = HW ? -0.5f : 0.0f;
  const float hx = HW ? fW - 0.5f : fW, hy = HW ? fH - 0.5f : fH, hz = HW ? fD - 0.5f : fD;
  float amin = 0.0f, amax = __builtin_inff();
  if (0.0f != rx) {
    const float reci = 1.0f / rx;
    const float a0 = (lo - s0) * reci, a1 = (hx - s0) * reci;
    amin = fminf(a0, a1);
    amax = fmaxf(a0, a1);
  }
  if (0.0f != ry) {
    const float reci = 1.0f / ry;
    const float a0 = (lo - s1) * reci, a1 = (hy - s1) * reci;
    amin = fmaxf(amin, fminf(a0, a1));
    amax = fminf(amax, fmaxf(a0, a1));
  }
  if (0.0f != rz) {
    const float reci = 1.0f / rz;
    const float a0 = (lo - s2) * reci, a1 = (hz - s2) * reci;
    amin = fmaxf(amin, fminf(a0, a1));
    amax = fminf(amax, fmaxf(a0, a1));
  }
  amin = fmaf(step, 0.5f, amin);

  // A ray needs neither the coordinate clamp nor the iteration cap when
  //  * every direction component is non-zero, so all three slabs clip the interval, and
  //    |source| < 2^18 and |alpha| < 2^19: the slab parameters carry a relative error of a
  //    few 2^-23 and each position one rounding, so every sample lies within 0.25 voxel
  //    of the clip box and its floor indexes an existing cell;
  //  * step >= 2^-20 |alpha|, i.e. at least 8 ulp of alpha: each alpha += step advances by
  //    at least 15/16 step, so the loop ends before 1.07 (amax - amin) / step + 1
  //    iterations (at most 2^21 * 1.07 + 1), under the cap of 1.25 (amax - amin) / step + 16
  //    and under its 2^25 ceiling.
  // The guarded march is used unless every lane of the wave qualifies, so the choice is
  // uniform and never changes a result.
  const float amag = fmaxf(fabsf(amin), fabsf(amax));
  const bool plain = rx != 0.0f && ry != 0.0f && rz != 0.0f &&
                     fmaxf(fabsf(s0), fmaxf(fabsf(s1), fabsf(s2))) < 0x1p18f && amag < 0x1p19f &&
                     amag * 0x1p-20f <= step;
  float pix;
  if constexpr (sizeof(TS) == 2 && !DIRECT)
    pix = __all(plain)
              ? march_ray_stream<TS, HW, false, IT>(v, Wp, HWp, fW, fH, fD, rx, ry, rz, s0, s1, s2, amin, amax, step)
              : march_ray_stream<TS, HW, true, IT>(v, Wp, HWp, fW, fH, fD, rx, ry, rz, s0, s1, s2, amin, amax, step);
  else
    pix = __all(plain)
              ? march_ray<TS, HW, false, IT, DIRECT>(v, Wp, HWp, fW, fH, fD, rx, ry, rz, s0, s1, s2, amin, amax, step)
              : march_ray<TS, HW, true, IT, DIRECT>(v, Wp, HWp, fW, fH, fD, rx, ry, rz, s0, s1, s2, amin, amax, step);
  const float a = rx * spx, bq = ry * spy, c = rz * spz;
  if (store) *out = pix * sqrtf(fmaf(c, c, fmaf(bq, bq, a * a)));
}

// One launch covers the views P.n of a range of each item's n_all views (the host splits
// the views so that a launch stays within the dispatch grid's 2^32 work-items); P.sino,
// P.inv_ar and P.src point at the range's first view of item 0.  cell_stride is the
// elements per batch item of P.cells (the volume's with DIRECT).
template <typename TS, bool HW, typename IT, bool DIRECT = false, typename TL = FwdTile16>
__global__ void __launch_bounds__(256)
cone_fwd_kernel(FwdProblem P, long long cell_stride, int n_all) {
  // blockIdx.z selects the batch item; items share shapes and have their own geometry
  const int b = blockIdx.z;
  P.cells = static_cast<const TS *>(P.cells) + (size_t)b * cell_stride;
  P.sino += (size_t)b * n_all * P.Hd * P.Wd;
  P.inv_ar += (size_t)b * n_all * 9;
  P.src += (size_t)b * n_all * 3;
  P.step += b;
  P.spacing += (size_t)b * 3;
  fwd_block<TS, HW, IT, DIRECT, TL>(P, blockIdx.x);
}

// Up to this many problems of different shapes run in one persistent launch.
constexpr int kMaxBatch = PYRONN_MAX_BATCHED_PROBLEMS;

struct FwdBatch {
  FwdProblem p[kMaxBatch];
  int count;
  long long total;
};

// Work claiming for the persistent kernels: one 64-bit counter per XCD (zeroed by
// reset_counters_kernel before each launch).  A block on XCD x (block id mod 8; the
// grid is a multiple of 8) claims the next item of that XCD's sequence with a global
// atomic; the claim order decides only which workgroup computes an item, never the
// arithmetic, so results stay bitwise identical to single launches.
__global__ void reset_counters_kernel(unsigned long long *__restrict__ counters) {
  if (threadIdx.x < kXcds) counters[threadIdx.x] = 0ull;
}

__device__ __forceinline__ long long claim(unsigned long long *counters, int xcd) {
  __shared__ long long s_item;
  if (threadIdx.x == 0) s_item = (long long)atomicAdd(counters + xcd, 1ull);
  __syncthreads();
  const long long k = s_item;
  __syncthreads();
  return k;
}

// Persistent forward projection over a work list of uneven problems.  Items keep
// the single-launch order of each problem (problems concatenated, each a multiple of
// 8 items); XCD x processes items x, x + 8, x + 16, ... in claim order, so the blocks
// resident on an XCD always work on consecutive items of its detector row bands.
// TL is the block shape (the problems' tiles_x / tiles_y are counted in it).
template <typename TS, bool HW, typename IT, typename TL>
__global__ void __launch_bounds__(256)
cone_fwd_persistent_kernel(FwdBatch B, unsigned long long *counters) {
  // one claim takes kFwdClaim consecutive items of the XCD's sequence (a 16 x 16 tile of
  // one view is only tens of microseconds of work, so this amortises the atomic and the
  // two barriers of a claim); the per-XCD order of items is unchanged
  constexpr int kFwdClaim = 4;
  const int xcd = (int)(blockIdx.x % kXcds);
  for (;;) {
    const long long t0 = claim(counters, xcd) * kFwdClaim;
    if (t0 * kXcds + xcd >= B.total) break;
    int k = 0;
    for (int j = 0; j < kFwdClaim; ++j) {
      const long long w = (t0 + j) * kXcds + xcd;
      if (w >= B.total) break;
      while (k + 1 < B.count && w >= B.p[k + 1].begin) ++k;
      fwd_block<TS, HW, IT, false, TL>(B.p[k], w - B.p[k].begin);
    }
  }
}

// The same work list as one workgroup per item in item order, from item item0 (a multiple
// of 8) on: the hardware deals workgroups round-robin over the XCDs in block order, so item
// mod 8 is the XCD, as in the single-problem grid, and no claim (atomic and two barriers)
// precedes a block.
template <typename TS, bool HW, typename IT, typename TL>
__global__ void __launch_bounds__(256) cone_fwd_items_kernel(FwdBatch B, long long item0) {
  const long long w = item0 + blockIdx.x;
  int k = 0;
  while (k + 1 < B.count && w >= B.p[k + 1].begin) ++k;
  fwd_block<TS, HW, IT, false, TL>(B.p[k], w - B.p[k].begin);
}

// =====================================================================================
// Detector cells for the backprojection: per view, cell (u, v), u in [-1, W_d],
// v in [-1, H_d], holds the bilinear neighbourhood
//   [s(u,v), s(u+1,v), s(u,v+1), s(u+1,v+1)]
// with zeros outside the detector, so one update is one 16-byte (f32) or 8-byte
// (bf16 / f16) load.
// =====================================================================================
template <typename TS>
__device__ __forceinline__ TS pix(const TS *__restrict__ s, int u, int v, int Wd, int Hd) {
  if (u < 0 || u >= Wd || v < 0 || v >= Hd) return zero_of<TS>();
  return s[(size_t)v * Wd + u];
}

// Lanes along a row of cells (x), blocks over runs of kDetCellRows cell rows (y) and
// images (z), so no lane divides.  Walking down its column, a lane keeps the detector
// row it read for the cell above and loads only the next row: two loads and one vector
// store per cell, every cell written by exactly one lane.
constexpr int kDetCellRows = 16;

template <typename TS>
__global__ void build_det_cells_kernel(const TS *__restrict__ sino, TS *__restrict__ cells,
                                       int Hd, int Wd, int images) {
  const int Wc = Wd + 2, Hc = Hd + 2;
  const int cu = blockIdx.x * blockDim.x + threadIdx.x;
  if (cu >= Wc) return;
  const int u = cu - 1;
  for (int img = blockIdx.z; img < images; img += gridDim.z) {
    const TS *s = sino + (size_t)img * Hd * Wd;
    for (int cv0 = blockIdx.y * kDetCellRows; cv0 < Hc; cv0 += gridDim.y * kDetCellRows) {
      const int cv1 = cv0 + kDetCellRows < Hc ? cv0 + kDetCellRows : Hc;
      TS *o = cells + (((size_t)img * Hc + cv0) * Wc + cu) * 4;
      // cell row cv covers detector rows v = cv - 1 and v + 1 = cv
      TS a = pix(s, u, cv0 - 1, Wd, Hd), b = pix(s, u + 1, cv0 - 1, Wd, Hd);
#pragma unroll 4
      for (int cv = cv0; cv < cv1; ++cv, o += (size_t)Wc * 4) {
        const TS c = pix(s, u, cv, Wd, Hd), d = pix(s, u + 1, cv, Wd, Hd);
        store_cell(o, a, b, c, d);
        a = c;
        b = d;
      }
    }
  }
}

// Bilinear sample, lerp x then y (cone_backprojector_3D_CudaKernel.cu.cc:55-82),
// zero outside the detector.
// FIDX: img points at cell (0, 0) (not at the cell array's start, cell (-1, -1)) and the
// host guarantees (H_d + 2)(W_d + 2) <= 2^24, so the cell offset fy W_c + fx is an integer
// of magnitude below 2^24 that one float FMA forms exactly (two instructions fewer than
// the integer form).
template <typename TS, bool HW, bool FIDX>
__device__ __forceinline__ float sample_det(const TS *__restrict__ img, int Wc, float u, float v,
                                            float fW, float fH) {
  if (HW) {  // tex2DLayered(u + 0.5, v + 0.5, n)
    u = (u + 0.5f) - 0.5f;
    v = (v + 0.5f) - 0.5f;
  }
  u = clamp_apron(u, fW);
  v = clamp_apron(v, fH);
  const float fx = floorf(u), fy = floorf(v);
  const float xd = u - fx, yd = v - fy;
  // signed 24-bit multiply (full rate); the host keeps W_d + 2 and H_d + 2 below 2^23
  const int c = FIDX ? (int)fmaf(fy, (float)Wc, fx) : __mul24((int)fy, Wc) + (int)fx + (Wc + 1);
  float p00, p01, p10, p11;
  if constexpr (sizeof(TS) == 4) {
    const float4 a = reinterpret_cast<const float4 *>(img)[c];
    p00 = a.x; p01 = a.y; p10 = a.z; p11 = a.w;
  } else {
    const uint2 a = reinterpret_cast<const uint2 *>(img)[c];
    unpack2(a.x, TS{}, p00, p01);
    unpack2(a.y, TS{}, p10, p11);
  }
  const float p0 = lerp_ref(p00, p01, xd);
  const float p1 = lerp_ref(p10, p11, xd);
  return lerp_ref(p0, p1, yd);
}

// =====================================================================================
// Backprojection.  A 256-lane block covers 16 x 16 voxels in (x, y); each lane
// accumulates ZPT consecutive z voxels, sharing the x/y part of P x between them
// (the products are formed in the reference's order m0 x, + m1 y, + m2 z, + m3, so
// sharing the first two terms changes nothing).  P rows are wave-uniform and come
// through scalar loads.  cone_backprojector_3D_CudaKernel.cu.cc:44-53 (map),
// :84-118 (kernel); texture variant cone_backprojector_3D_CudaKernel_hardware_interp.cu.cc:55-87.
// =====================================================================================
struct BpProblem {
  const void *cells;     // detector cells, [n][(H_d + 2) (W_d + 2)] x 4
  float *vol;            // [D][H][W]
  const float *P;        // [n][12]
  const float *mult;     // [1]
  const float *spacing;  // [3] (z, y, x)
  const float *origin;   // [3] (z, y, x)
  int W, H, D, n, Hd, Wd, tiles_x, tiles_y, zblocks;
  long long begin;
  int flags;             // kBpContinue: start from the running sums in vol; kBpScale: write sum * mult
  int z0, z1;            // the z slices [z0, z1) of the volume this launch computes
};
constexpr int kBpContinue = 1, kBpScale = 2;
// XCD-compact tiles: the (x, y) tiles of a z block form a 4 x 2 grid of regions and XCD k
// (work item mod 8) takes region k, row by row, so the blocks resident on an XCD are
// neighbours whose detector footprints overlap in its L2 (f32 72.6 against 76.1 ms with
// consecutive tiles dealt round-robin over the XCDs, on the 512^3 / 720-view
// backprojection; 8 x 1, 2 x 4 and 1 x 8 region grids were in between or slower).  A z
// block has 8 equal regions, padded; padding items do nothing.  Work items of a z block
// are numbered from a multiple of 8, so that item mod 8 is the XCD the item runs on.
constexpr int kBpRegionsX = 4, kBpRegionsY = 2;
static_assert(kBpRegionsX * kBpRegionsY == kXcds, "one region per XCD");
__host__ __device__ inline int bp_region_w(int tiles_x) { return (tiles_x + kBpRegionsX - 1) / kBpRegionsX; }
__host__ __device__ inline int bp_region_h(int tiles_y) { return (tiles_y + kBpRegionsY - 1) / kBpRegionsY; }
// Work items per z block (the padded tile count).
__host__ __device__ inline int bp_items_per_z(int tiles_x, int tiles_y) {
  return kXcds * bp_region_w(tiles_x) * bp_region_h(tiles_y);
}
// Tile (tx, ty) of item r of a z block; false for a padding item.
__device__ __forceinline__ bool bp_region_tile(int r, int tiles_x, int tiles_y, int &tx, int &ty) {
  const int rw = bp_region_w(tiles_x), rh = bp_region_h(tiles_y);
  const int xcd = r % kXcds, j = r / kXcds;
  tx = (xcd % kBpRegionsX) * rw + j % rw;
  ty = (xcd / kBpRegionsX) * rh + j / rw;
  return tx < tiles_x && ty < tiles_y;
}

template <typename TS, bool HW, int ZPT, bool FIDX>
__device__ __forceinline__ void bp_block(const BpProblem &Q, int tx, int ty, int zb) {
  // Each group of four lanes (what the texture-address unit takes per cycle; it spends a
  // cycle per distinct 128-byte line in the group) is a 2 x 2 voxel square, whose detector
  // footprint is narrower than that of four voxels in a row for most view angles.  Lane
  // bits: x0 y0 x1 x2 x3 y1, wave = y2 y3.
  const int l = threadIdx.x & 63;
  const int i = tx * 16 + ((l & 1) | (((l >> 2) & 7) << 1));
  const int j = ty * 16 + (((l >> 1) & 1) | ((l >> 5) << 1) | ((threadIdx.x >> 6) << 2));
  const int k0 = Q.z0 + zb * ZPT;
  const int W = Q.W, H = Q.H, z1 = Q.z1, n = Q.n, Hd = Q.Hd, Wd = Q.Wd;
  if (i >= W || j >= H) return;

  const float sx = Q.spacing[2], sy = Q.spacing[1], sz = Q.spacing[0];
  const float ox = Q.origin[2], oy = Q.origin[1], oz = Q.origin[0];
  const float x = fmaf((float)i, sx, ox), y = fmaf((float)j, sy, oy);
  float z[ZPT], acc[ZPT];
#pragma unroll
  for (int q = 0; q < ZPT; ++q) {
    z[q] = fmaf((float)(k0 + q), sz, oz);
    // a running sum continued from an earlier range of views is the same float the
    // accumulation would hold here in one pass, so chunked calls reproduce it exactly
    acc[q] = ((Q.flags & kBpContinue) && k0 + q < z1) ? Q.vol[((size_t)(k0 + q) * H + j) * W + i] : 0.0f;
  }
  const int Wc = Wd + 2;
  const size_t img_stride = (size_t)(Hd + 2) * Wc * 4;
  const TS *img = static_cast<const TS *>(Q.cells) + (FIDX ? (size_t)(Wc + 1) * 4 : 0);
  const float *Pm = Q.P;
  const float fW = (float)Wd, fH = (float)Hd;

  for (int view = 0; view < n; ++view, img += img_stride) {
    const float *m = Pm + view * 12;
    const float tu = fmaf(m[1], y, m[0] * x);
    const float tv = fmaf(m[5], y, m[4] * x);
    const float tw = fmaf(m[9], y, m[8] * x);
    // Detector column axis and depth perpendicular to z (m[2] = m[10] = 0, as for any
    // orbit about z with an untilted detector): the reference's per-voxel w and u are
    // then the same floats for all z of the lane.  fmaf(0, z, t) + c does not depend on
    // z unless it is a signed zero; w0 != 0 excludes that case for w, and u = +-0 samples
    // the detector identically (the fraction u - floor(u) is +0 either way), so one
    // reciprocal serves the lane's ZPT voxels, bit for bit.
    if (m[2] == 0.0f && m[10] == 0.0f) {
      const float w0 = fmaf(m[10], z[0], tw) + m[11];
      if (w0 != 0.0f) {
        const float ww = rcp_rn(w0);
        const float uu = (fmaf(m[2], z[0], tu) + m[3]) * ww;
#pragma unroll
        for (int q = 0; q < ZPT; ++q) {
          const float vv = (fmaf(m[6], z[q], tv) + m[7]) * ww;
          const float f = sample_det<TS, HW, FIDX>(img, Wc, uu, vv, fW, fH);
          acc[q] = fmaf(f * ww, ww, acc[q]);
        }
        continue;
      }
    }
#pragma unroll
    for (int q = 0; q < ZPT; ++q) {
      const float ww = rcp_rn(fmaf(m[10], z[q], tw) + m[11]);  // correctly rounded, as the reference's division
      const float uu = (fmaf(m[2], z[q], tu) + m[3]) * ww;
      const float vv = (fmaf(m[6], z[q], tv) + m[7]) * ww;
      const float f = sample_det<TS, HW, FIDX>(img, Wc, uu, vv, fW, fH);
      acc[q] = fmaf(f * ww, ww, acc[q]);
    }
  }
  const bool scale = (Q.flags & kBpScale) != 0;
  const float mu = scale ? Q.mult[0] : 1.0f;
#pragma unroll
  for (int q = 0; q < ZPT; ++q) {
    if (k0 + q < z1) Q.vol[((size_t)(k0 + q) * H + j) * W + i] = scale ? acc[q] * mu : acc[q];
  }
}

// Block (x, z) of a launch: item item0 + x of z block z mod zblocks of batch item
// z / zblocks (item0 is a multiple of 8, so that block x runs on the XCD of its item).
template <typename TS, bool HW, int ZPT, bool FIDX>
__global__ void __launch_bounds__(256, 5) cone_bp_kernel(BpProblem Q, size_t cells_stride, int item0) {
  const int b = blockIdx.z / Q.zblocks;
  const int zb = blockIdx.z - b * Q.zblocks;
  Q.cells = static_cast<const TS *>(Q.cells) + (size_t)b * cells_stride;
  Q.vol += (size_t)b * Q.D * Q.H * Q.W;
  Q.P += (size_t)b * Q.n * 12;
  Q.mult += b;
  Q.spacing += (size_t)b * 3;
  Q.origin += (size_t)b * 3;
  int tx, ty;
  if (!bp_region_tile(item0 + (int)blockIdx.x, Q.tiles_x, Q.tiles_y, tx, ty)) return;
  bp_block<TS, HW, ZPT, FIDX>(Q, tx, ty, zb);
}

struct BpBatch {
  BpProblem p[kMaxBatch];
  int count;
  long long total;
};

// Persistent backprojection over uneven problems; item -> (problem, z block, tile of the
// z block in region order).  XCD x takes items x, x + 8, ... in claim order, so each XCD
// walks its own region of every z block, as in the single-problem grid.
template <typename TS, bool HW, int ZPT, bool FIDX>
__global__ void __launch_bounds__(256)
cone_bp_persistent_kernel(BpBatch B, unsigned long long *counters) {
  const int xcd = (int)(blockIdx.x % kXcds);
  for (;;) {
    const long long w = claim(counters, xcd) * kXcds + xcd;
    if (w >= B.total) break;
    int k = 0;
    while (k + 1 < B.count && w >= B.p[k + 1].begin) ++k;
    const BpProblem &Q = B.p[k];
    const long long q = w - Q.begin;
    const int per_z = bp_items_per_z(Q.tiles_x, Q.tiles_y);
    const int zb = (int)(q / per_z);
    int tx, ty;
    if (bp_region_tile((int)(q - (long long)zb * per_z), Q.tiles_x, Q.tiles_y, tx, ty))
      bp_block<TS, HW, ZPT, FIDX>(Q, tx, ty, zb);
  }
}

// The same work list as one workgroup per item in item order (no claiming): the hardware
// dispatches workgroups round-robin over the XCDs in block order, so item mod 8 is the XCD,
// as in the single-problem grid.  The item decode is wave-uniform (scalar), so the kernel keeps
// the single kernel's register budget; the persistent kernel holds its claim loop
// around the block and needs more.
template <typename TS, bool HW, int ZPT, bool FIDX>
__global__ void __launch_bounds__(256, 5) cone_bp_items_kernel(BpBatch B) {
  const long long w = blockIdx.x;
  int k = 0;
  while (k + 1 < B.count && w >= B.p[k + 1].begin) ++k;
  const BpProblem &Q = B.p[k];
  const int q = (int)(w - Q.begin);
  const int per_z = bp_items_per_z(Q.tiles_x, Q.tiles_y);
  const int zb = q / per_z;
  int tx, ty;
  if (!bp_region_tile(q - zb * per_z, Q.tiles_x, Q.tiles_y, tx, ty)) return;
  bp_block<TS, HW, ZPT, FIDX>(Q, tx, ty, zb);
}

// =====================================================================================
// Host side
// =====================================================================================
static size_t cell_elems_dt(int dtype, int D, int H, int W) {
  return dtype == PYRONN_F32 ? cell_elems<float>(D, H, W) : cell_elems<bf16s>(D, H, W);
}

// Volumes the forward accepts: every side below 2^24 - 3, so that each cell coordinate and
// its floor are exact floats (the reference indexes with uint32 and so takes volumes below
// 2^32 voxels; all of those with sides below 2^24 - 3 are accepted here, and larger ones).
static bool fwd_volume_ok(int D, int H, int W) {
  const int lim = (1 << 24) - 3;
  return D < lim && H < lim && W < lim;
}
// Whether the forward may use the 32-bit cell index (sample_cells, IT = int): fewer than 2^31
// cells, and W + 2 and the (y, z) plane of cells below 2^23 for the 24-bit multiply and the
// exact float FMA of the sampler.  Larger volumes take the 64-bit index.
static bool fwd_small_index(int D, int H, int W) {
  const long long cells = (long long)(W + 2) * (H + 2) * (D + CellLayout<float>::kZExtra);
  return cells < (long long)INT32_MAX && W + 2 < (1 << 23) &&
         (long long)(H + 2) * (D + CellLayout<float>::kZExtra) < (1LL << 23);
}
// Whether the 16-bit forward may address its 8-byte cells by 32-bit byte offsets (locate16,
// IT = unsigned): 8 (H + 2) Dc below 2^23 for the exact float and the 24-bit multiply, and
// the cell array within 0xB5000000 bytes so that offset + 0x4B000000 stays below 2^32.
static bool fwd_byte_offsets(int D, int H, int W) {
  const long long plane = (long long)(H + 2) * (D + CellLayout<float>::kZExtra);
  return W + 2 < (1 << 23) && 8 * plane < (1LL << 23) && 8 * (W + 2) * plane <= 0xB5000000LL;
}
static size_t det_cells(int n, int Hd, int Wd) { return (size_t)n * (Hd + 2) * (Wd + 2); }
// Whether the backprojection may form detector cell offsets with one float FMA (exact below 2^24).
static bool det_fidx(int Hd, int Wd) { return (long long)(Hd + 2) * (Wd + 2) <= (1LL << 24); }

// A dispatch's grid size is a 32-bit count of work-items, so a 1-D grid of 256-lane blocks
// holds at most 2^24 - 1 blocks; grid y and z hold at most 65535 blocks here.
constexpr long long kMaxGridX = (1LL << 24) - 1;
constexpr int kMaxGridYZ = 65535;

// Blocks of one forward launch per batch item: 8 XCD lanes x tile-row bands x views x tiles
// (tile shape `tile`, kFwdTiles).
static long long fwd_blocks(int n, int Hd, int Wd, int tile = 0) {
  const int tc = fwd_tile_cols(tile), tr = fwd_tile_rows(tile);
  const long long tiles_x = (Wd + tc - 1) / tc, tiles_y = (Hd + tr - 1) / tr;
  return (long long)kXcds * ((tiles_y + kXcds - 1) / kXcds) * n * tiles_x;
}
// Views per forward launch (0: one view alone exceeds a launch).
static int fwd_views_per_launch(int n, int Hd, int Wd, int tile = 0) {
  const long long per = kMaxGridX / fwd_blocks(1, Hd, Wd, tile);
  return (int)(per < n ? per : n);
}

// Tile shape (kFwdTiles) of the single-problem forward with a cell copy: 32 x 8 pixels of
// 8 x 8 waves for f32 cells, and for 16-bit cells of volumes up to 2^25 voxels; 16 x 16 of
// 4 x 16 waves for larger 16-bit volumes.  Measured per problem (phantom volumes, 16 x 16 /
// 4 x 16 against 32 x 8 / 8 x 8): f32 equal on the 512^3 / 720-view / 1024 x 768 workload
// (372.7 / 372.9 ms) and 3-11% faster on the others (256^3 / 360 views / 400^2: 27.8 / 24.8
// ms); bf16 faster up to 320^3 (43.8 / 41.0 ms), equal at 384^3 and slower beyond (512^3:
// 213.2 / 237.6 ms).  The shape changes which lane computes which ray, never a result.
template <typename TS> static int fwd_single_tile(int D, int H, int W) {
  return (sizeof(TS) == 4 || (long long)D * H * W <= (1LL << 25)) ? 2 : 0;
}

template <typename TS>
static void launch_cells(const void *vol, void *cells, int B, int D, int H, int W, hipStream_t stream) {
  const long long tiles = (long long)((W + 2 + kCellTileX - 1) / kCellTileX) *
                          ((D + CellLayout<TS>::kZExtra + kCellTileZ - 1) / kCellTileZ) * (H + 2) * B;
  const long long blocks = tiles < 65536 * 8 ? tiles : 65536 * 8;
  build_cells_kernel<TS><<<dim3((unsigned)blocks), dim3(256), 0, stream>>>((const TS *)vol, (TS *)cells, W, H,
                                                                           D, B);
}

template <typename TS, typename IT>
static void launch_fwd_it(bool hw, int tile, const FwdProblem &P, dim3 grid, long long cs, int n_all,
                          hipStream_t stream) {
  using T2 = FwdTile<32, 8, 8>;
  if (tile == 2) {
    if (hw)
      cone_fwd_kernel<TS, true, IT, false, T2><<<grid, dim3(256), 0, stream>>>(P, cs, n_all);
    else
      cone_fwd_kernel<TS, false, IT, false, T2><<<grid, dim3(256), 0, stream>>>(P, cs, n_all);
  } else if (hw) {
    cone_fwd_kernel<TS, true, IT><<<grid, dim3(256), 0, stream>>>(P, cs, n_all);
  } else {
    cone_fwd_kernel<TS, false, IT><<<grid, dim3(256), 0, stream>>>(P, cs, n_all);
  }
}

template <typename TS>
static void launch_fwd(bool hw, const void *cells, float *sino, const float *inv_ar,
                       const float *src, const float *step, const float *spacing, int B, int D,
                       int H, int W, int n, int Hd, int Wd, hipStream_t stream) {
  const long long cs = (long long)cell_elems<TS>(D, H, W);
  const bool small = fwd_small_index(D, H, W);
  const int tile = fwd_single_tile<TS>(D, H, W);
  const int tc = fwd_tile_cols(tile), tr = fwd_tile_rows(tile);
  const int per = fwd_views_per_launch(n, Hd, Wd, tile);
  for (int v0 = 0; v0 < n; v0 += per) {
    const int k = n - v0 < per ? n - v0 : per;
    FwdProblem P{cells, sino + (size_t)v0 * Hd * Wd, inv_ar + (size_t)v0 * 9, src + (size_t)v0 * 3, step,
                 spacing, W, H, D, k, Hd, Wd, (Wd + tc - 1) / tc, (Hd + tr - 1) / tr, 0};
    dim3 grid((unsigned)fwd_blocks(k, Hd, Wd, tile), 1, B);
    if constexpr (sizeof(TS) == 2) {
      if (fwd_byte_offsets(D, H, W)) {
        launch_fwd_it<TS, unsigned>(hw, tile, P, grid, cs, n, stream);
        continue;
      }
    }
    if (small)
      launch_fwd_it<TS, int>(hw, tile, P, grid, cs, n, stream);
    else
      launch_fwd_it<TS, long long>(hw, tile, P, grid, cs, n, stream);
  }
}

// The forward that reads the volume directly (sample_direct, no cell copy) with a 64-bit
// volume index; the same views per launch and block order as launch_fwd.
template <typename TS>
static void launch_fwd_direct(bool hw, const void *vol, float *sino, const float *inv_ar,
                              const float *src, const float *step, const float *spacing, int B, int D,
                              int H, int W, int n, int Hd, int Wd, hipStream_t stream) {
  const long long vs = (long long)D * H * W;
  const int per = fwd_views_per_launch(n, Hd, Wd);
  for (int v0 = 0; v0 < n; v0 += per) {
    const int k = n - v0 < per ? n - v0 : per;
    FwdProblem P{vol, sino + (size_t)v0 * Hd * Wd, inv_ar + (size_t)v0 * 9, src + (size_t)v0 * 3, step,
                 spacing, W, H, D, k, Hd, Wd, (Wd + 15) / 16, (Hd + 15) / 16, 0};
    dim3 grid((unsigned)fwd_blocks(k, Hd, Wd), 1, B);
    if (hw)
      cone_fwd_kernel<TS, true, long long, true><<<grid, dim3(256), 0, stream>>>(P, vs, n);
    else
      cone_fwd_kernel<TS, false, long long, true><<<grid, dim3(256), 0, stream>>>(P, vs, n);
  }
}

// z voxels per lane: 8 for f32 and bf16 detector cells, 4 for f16 ones (measured on the
// 512^3 / 720-view / 1024x768 workload: f32 63.0 against 70.9 ms with 4, bf16 53.7 against
// 57.7 ms, f16 56.9 ms with 4 against 125.9 ms with 8).
template <typename TS> constexpr int bp_zpt() { return std::is_same<TS, _Float16>::value ? 4 : 8; }

template <typename TS>
static void launch_det_cells(const void *sino, void *cells, int images, int Hd, int Wd,
                             hipStream_t stream) {
  const int Hc = Hd + 2;
  const int runs = (Hc + kDetCellRows - 1) / kDetCellRows;
  const dim3 grid((unsigned)((Wd + 2 + 255) / 256), (unsigned)(runs < kMaxGridYZ ? runs : kMaxGridYZ),
                  (unsigned)(images < kMaxGridYZ ? images : kMaxGridYZ));
  build_det_cells_kernel<TS><<<grid, dim3(256), 0, stream>>>((const TS *)sino, (TS *)cells, Hd, Wd,
                                                             images);
}

// Whether the item numbers of one backprojection z block of a W x H volume fit an int
// (a launch takes at most kMaxGridX of them; larger planes take several launches).
static bool bp_plane_fits(int H, int W) {
  return (long long)kXcds * bp_region_w((W + 15) / 16) * bp_region_h((H + 15) / 16) < (long long)INT32_MAX;
}

template <typename TS>
static void launch_bp(bool hw, const void *cells, float *vol, const float *P, const float *mult,
                      const float *spacing, const float *origin, int B, int D, int H, int W, int n,
                      int Hd, int Wd, hipStream_t stream, int flags = kBpScale, int z0 = 0,
                      int z1 = -1) {
  constexpr int Z = bp_zpt<TS>();
  if (z1 < 0) z1 = D;
  const int zblocks = (z1 - z0 + Z - 1) / Z;
  // grid z = batch items x z blocks, at most 65535 per launch: split the z blocks; grid x =
  // the items of a z block, at most kMaxGridX (rounded down to a multiple of 8) per launch
  const int per = kMaxGridYZ / B;
  const size_t cs = det_cells(n, Hd, Wd) * 4;
  const int items = bp_items_per_z((W + 15) / 16, (H + 15) / 16);
  const int max_x = (int)(kMaxGridX / kXcds * kXcds);
  for (int zb0 = 0; zb0 < zblocks; zb0 += per) {
    const int nz = zblocks - zb0 < per ? zblocks - zb0 : per;
    BpProblem Q{cells, vol, P, mult, spacing, origin, W, H, D, n, Hd, Wd,
                (W + 15) / 16, (H + 15) / 16, nz, 0, flags, z0 + zb0 * Z, z1};
    for (int i0 = 0; i0 < items; i0 += max_x) {
      dim3 grid((unsigned)(items - i0 < max_x ? items - i0 : max_x), 1, nz * B);
      if (det_fidx(Hd, Wd)) {
        if (hw)
          cone_bp_kernel<TS, true, Z, true><<<grid, dim3(256), 0, stream>>>(Q, cs, i0);
        else
          cone_bp_kernel<TS, false, Z, true><<<grid, dim3(256), 0, stream>>>(Q, cs, i0);
      } else {
        if (hw)
          cone_bp_kernel<TS, true, Z, false><<<grid, dim3(256), 0, stream>>>(Q, cs, i0);
        else
          cone_bp_kernel<TS, false, Z, false><<<grid, dim3(256), 0, stream>>>(Q, cs, i0);
      }
    }
  }
}

}  // namespace pyronn

using namespace pyronn;

extern "C" int pyronn_cone_geometry_setup(const float *projection_matrices,
                                          const float *volume_origin, const float *volume_spacing,
                                          int batch, int n_views, float *inv_ar, float *src,
                                          pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(batch > 0 && n_views > 0, "cone_geometry_setup: batch=%d n_views=%d", batch,
                   n_views);
  PYRONN_CHECK_ARG(projection_matrices && volume_origin && volume_spacing && inv_ar && src,
                   "cone_geometry_setup: null pointer");
  const int total = batch * n_views;
  cone_setup_kernel<<<dim3((total + 127) / 128), dim3(128), 0, (hipStream_t)stream>>>(
      projection_matrices, volume_origin, volume_spacing, batch, n_views, inv_ar, src);
  PYRONN_CHECK_LAUNCH("cone_setup_kernel");
  return PYRONN_OK;
}

static size_t fwd_geom_bytes(int batch, int n) { return align_up((size_t)batch * n * 12 * 4, 256); }

extern "C" size_t pyronn_cone_projection3d_workspace_size(int batch, int D, int H, int W,
                                                          int n_views, int volume_dtype) {
  if (batch <= 0 || D <= 0 || H <= 0 || W <= 0 || n_views <= 0 || !dtype_ok(volume_dtype)) return 0;
  return fwd_geom_bytes(batch, n_views) +
         align_up((size_t)batch * cell_elems_dt(volume_dtype, D, H, W) * dtype_bytes(volume_dtype), 256);
}

// Argument checks shared by the forward entry points.
static int fwd_check_args(const char *name, const void *volume, int volume_dtype, int batch, int D, int H,
                          int W, const float *volume_origin, const float *volume_spacing,
                          const float *projection_matrices, const float *step_size, int n_views,
                          int det_h, int det_w, const float *sinogram) {
  PYRONN_CHECK_ARG(volume && volume_origin && volume_spacing && projection_matrices && step_size && sinogram,
                   "%s: null pointer", name);
  PYRONN_CHECK_ARG(dtype_ok(volume_dtype), "%s: bad volume dtype %d", name, volume_dtype);
  PYRONN_CHECK_ARG(batch > 0 && batch <= kMaxGridYZ && D > 0 && H > 0 && W > 0 && n_views > 0 &&
                       det_h > 0 && det_w > 0,
                   "%s: bad shape B=%d D=%d H=%d W=%d n=%d det=%dx%d", name, batch, D, H, W, n_views,
                   det_h, det_w);
  PYRONN_CHECK_ARG((long long)batch * n_views < (long long)INT32_MAX, "%s: batch x views %lld too large",
                   name, (long long)batch * n_views);
  if (!fwd_volume_ok(D, H, W)) {
    set_error("%s: volume %dx%dx%d: a side of 2^24 - 3 voxels or more", name, D, H, W);
    return PYRONN_ERR_TOO_LARGE;
  }
  if (fwd_views_per_launch(n_views, det_h, det_w, 0) < 1 || fwd_views_per_launch(n_views, det_h, det_w, 2) < 1) {
    set_error("%s: detector %dx%d exceeds one launch", name, det_h, det_w);
    return PYRONN_ERR_TOO_LARGE;
  }
  return PYRONN_OK;
}

extern "C" int pyronn_cone_projection3d(const void *volume, int volume_dtype, int batch, int D,
                                        int H, int W, const float *volume_origin,
                                        const float *volume_spacing,
                                        const float *projection_matrices, const float *step_size,
                                        const float *projection_multiplier, int n_views, int det_h,
                                        int det_w, int hardware_interp, float *sinogram,
                                        void *workspace, size_t workspace_bytes,
                                        pyronn_stream_t stream) {
  (void)projection_multiplier;  // read but unused by the reference (cone_projector_3D_OPKernel.cc:213-215)
  clear_error();
  int rc = fwd_check_args("cone_projection3d", volume, volume_dtype, batch, D, H, W, volume_origin,
                          volume_spacing, projection_matrices, step_size, n_views, det_h, det_w, sinogram);
  if (rc != PYRONN_OK) return rc;
  PYRONN_CHECK_ARG(workspace != nullptr, "cone_projection3d: null pointer");
  const size_t need = pyronn_cone_projection3d_workspace_size(batch, D, H, W, n_views, volume_dtype);
  if (workspace_bytes < need) {
    set_error("cone_projection3d: workspace %zu bytes < required %zu", workspace_bytes, need);
    return PYRONN_ERR_WORKSPACE;
  }
  PYRONN_CHECK_ARG(((uintptr_t)workspace & 15) == 0, "cone_projection3d: workspace must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  char *ws = (char *)workspace;
  float *inv_ar = (float *)ws;
  float *srcp = inv_ar + (size_t)batch * n_views * 9;
  void *cells = ws + fwd_geom_bytes(batch, n_views);

  rc = pyronn_cone_geometry_setup(projection_matrices, volume_origin, volume_spacing, batch, n_views,
                                  inv_ar, srcp, stream);
  if (rc != PYRONN_OK) return rc;
  const bool hw = hardware_interp != 0;
  if (volume_dtype == PYRONN_F32) {
    launch_cells<float>(volume, cells, batch, D, H, W, s);
    launch_fwd<float>(hw, cells, sinogram, inv_ar, srcp, step_size, volume_spacing, batch, D, H, W,
                      n_views, det_h, det_w, s);
  } else if (volume_dtype == PYRONN_BF16) {
    launch_cells<bf16s>(volume, cells, batch, D, H, W, s);
    launch_fwd<bf16s>(hw, cells, sinogram, inv_ar, srcp, step_size, volume_spacing, batch, D, H, W,
                      n_views, det_h, det_w, s);
  } else {
    launch_cells<_Float16>(volume, cells, batch, D, H, W, s);
    launch_fwd<_Float16>(hw, cells, sinogram, inv_ar, srcp, step_size, volume_spacing, batch, D, H,
                         W, n_views, det_h, det_w, s);
  }
  PYRONN_CHECK_LAUNCH("cone_fwd_kernel");
  return PYRONN_OK;
}

// =====================================================================================
// The forward projection that allocates nothing (for callers that cannot allocate device
// memory: code under HIP graph capture, serving loops, harnesses that forbid it).  Its
// only scratch is the per-view geometry, in a file-scope device array of fixed size; the
// volume is sampled directly (sample_direct) instead of through a cell copy.
// =====================================================================================
namespace pyronn {
constexpr int kNoAllocMaxViews = 1 << 16;  // batch x views of one call
constexpr int kNoAllocMaxDevices = 64;
// inv_ar [batch][n][9], then src [batch][n][3]: 3 MiB, held by each device the library
// runs on until the process ends.
static __device__ __align__(256) float scratch[(size_t)kNoAllocMaxViews * 12];

// The device address of `scratch` on device `dev` (the current device), looked up once per
// device.  In host code the array's name is only a host-side stand-in; kernels get this
// address.
static float *scratch_address(int dev, int &rc) {
  static std::mutex mu;
  static float *addr[kNoAllocMaxDevices] = {};
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      rc = hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "cone_projection3d_noalloc: hipGetSymbolAddress");
      return nullptr;
    }
    addr[dev] = static_cast<float *>(p);
  }
  return addr[dev];
}
}  // namespace pyronn

extern "C" int pyronn_cone_projection3d_noalloc(const void *volume, int volume_dtype, int batch, int D,
                                                int H, int W, const float *volume_origin,
                                                const float *volume_spacing,
                                                const float *projection_matrices,
                                                const float *step_size,
                                                const float *projection_multiplier, int n_views,
                                                int det_h, int det_w, int hardware_interp,
                                                float *sinogram, pyronn_stream_t stream) {
  (void)projection_multiplier;  // read but unused by the reference (cone_projector_3D_OPKernel.cc:213-215)
  clear_error();
  int rc = fwd_check_args("cone_projection3d_noalloc", volume, volume_dtype, batch, D, H, W, volume_origin,
                          volume_spacing, projection_matrices, step_size, n_views, det_h, det_w, sinogram);
  if (rc != PYRONN_OK) return rc;
  if ((long long)batch * n_views > kNoAllocMaxViews) {
    set_error("cone_projection3d_noalloc: batch x views = %lld exceeds the scratch limit of %d",
              (long long)batch * n_views, kNoAllocMaxViews);
    return PYRONN_ERR_TOO_LARGE;
  }
  int dev = 0;
  const hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "cone_projection3d_noalloc: hipGetDevice");
  PYRONN_CHECK_ARG(dev >= 0 && dev < kNoAllocMaxDevices, "cone_projection3d_noalloc: device %d not supported", dev);
  float *inv_ar = scratch_address(dev, rc);
  if (inv_ar == nullptr) return rc;
  float *srcp = inv_ar + (size_t)batch * n_views * 9;
  rc = pyronn_cone_geometry_setup(projection_matrices, volume_origin, volume_spacing, batch, n_views,
                                  inv_ar, srcp, stream);
  if (rc != PYRONN_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const bool hw = hardware_interp != 0;
  if (volume_dtype == PYRONN_F32)
    launch_fwd_direct<float>(hw, volume, sinogram, inv_ar, srcp, step_size, volume_spacing, batch, D, H, W,
                             n_views, det_h, det_w, s);
  else if (volume_dtype == PYRONN_BF16)
    launch_fwd_direct<bf16s>(hw, volume, sinogram, inv_ar, srcp, step_size, volume_spacing, batch, D, H, W,
                             n_views, det_h, det_w, s);
  else
    launch_fwd_direct<_Float16>(hw, volume, sinogram, inv_ar, srcp, step_size, volume_spacing, batch, D, H,
                                W, n_views, det_h, det_w, s);
  PYRONN_CHECK_LAUNCH("cone_fwd_kernel (direct)");
  return PYRONN_OK;
}

// Shape checks shared by the backprojection entry points: the detector cell offset of the
// sampler (24-bit multiplies, below 2^31 per image), one z block of the volume per launch
// row, and batch x views images indexed with int.
static int bp_check_shape(const char *name, int batch, int D, int H, int W, int n, int det_h, int det_w) {
  PYRONN_CHECK_ARG(batch > 0 && batch <= kMaxGridYZ && D > 0 && H > 0 && W > 0 && n > 0 && det_h > 0 &&
                       det_w > 0,
                   "%s: bad shape B=%d D=%d H=%d W=%d n=%d det=%dx%d", name, batch, D, H, W, n, det_h, det_w);
  PYRONN_CHECK_ARG((long long)batch * n < (long long)INT32_MAX, "%s: batch x views %lld too large", name,
                   (long long)batch * n);
  if ((size_t)(det_h + 2) * (det_w + 2) >= (size_t)INT32_MAX / 4 || det_w + 2 >= (1 << 23) ||
      det_h + 2 >= (1 << 23)) {
    set_error("%s: detector %dx%d too large", name, det_h, det_w);
    return PYRONN_ERR_TOO_LARGE;
  }
  if (!bp_plane_fits(H, W)) {
    set_error("%s: volume plane %dx%d too large for one launch", name, H, W);
    return PYRONN_ERR_TOO_LARGE;
  }
  return PYRONN_OK;
}

extern "C" size_t pyronn_cone_backprojection3d_workspace_size(int batch, int n_views, int det_h,
                                                              int det_w, int sinogram_dtype) {
  if (batch <= 0 || n_views <= 0 || det_h <= 0 || det_w <= 0 || !dtype_ok(sinogram_dtype)) return 0;
  return align_up((size_t)batch * det_cells(n_views, det_h, det_w) * 4 * dtype_bytes(sinogram_dtype),
                  256);
}

extern "C" int pyronn_cone_backprojection3d(const void *sinogram, int sinogram_dtype, int batch,
                                            int n_views, int det_h, int det_w,
                                            const float *volume_origin,
                                            const float *volume_spacing,
                                            const float *projection_matrices,
                                            const float *step_size,
                                            const float *projection_multiplier, int D, int H,
                                            int W, int hardware_interp, float *volume,
                                            void *workspace, size_t workspace_bytes,
                                            pyronn_stream_t stream) {
  (void)step_size;  // read but unused by the reference (cone_backprojector_3D_OPKernel.cc:99-101)
  clear_error();
  PYRONN_CHECK_ARG(sinogram && volume_origin && volume_spacing && projection_matrices &&
                       projection_multiplier && volume && workspace,
                   "cone_backprojection3d: null pointer");
  PYRONN_CHECK_ARG(dtype_ok(sinogram_dtype), "cone_backprojection3d: bad sinogram dtype %d",
                   sinogram_dtype);
  const int rc = bp_check_shape("cone_backprojection3d", batch, D, H, W, n_views, det_h, det_w);
  if (rc != PYRONN_OK) return rc;
  const size_t need =
      pyronn_cone_backprojection3d_workspace_size(batch, n_views, det_h, det_w, sinogram_dtype);
  if (workspace_bytes < need) {
    set_error("cone_backprojection3d: workspace %zu bytes < required %zu", workspace_bytes, need);
    return PYRONN_ERR_WORKSPACE;
  }
  PYRONN_CHECK_ARG(((uintptr_t)workspace & 15) == 0, "cone_backprojection3d: workspace must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const bool hw = hardware_interp != 0;
  const int images = batch * n_views;
  if (sinogram_dtype == PYRONN_F32) {
    launch_det_cells<float>(sinogram, workspace, images, det_h, det_w, s);
    launch_bp<float>(hw, workspace, volume, projection_matrices, projection_multiplier,
                     volume_spacing, volume_origin, batch, D, H, W, n_views, det_h, det_w, s);
  } else if (sinogram_dtype == PYRONN_BF16) {
    launch_det_cells<bf16s>(sinogram, workspace, images, det_h, det_w, s);
    launch_bp<bf16s>(hw, workspace, volume, projection_matrices, projection_multiplier,
                     volume_spacing, volume_origin, batch, D, H, W, n_views, det_h, det_w, s);
  } else {
    launch_det_cells<_Float16>(sinogram, workspace, images, det_h, det_w, s);
    launch_bp<_Float16>(hw, workspace, volume, projection_matrices, projection_multiplier,
                        volume_spacing, volume_origin, batch, D, H, W, n_views, det_h, det_w, s);
  }
  PYRONN_CHECK_LAUNCH("cone_bp_kernel");
  return PYRONN_OK;
}

extern "C" int pyronn_cone_backprojection3d_views(
    const void *sinogram, int sinogram_dtype, int batch, int n_views, int det_h, int det_w,
    const float *volume_origin, const float *volume_spacing, const float *projection_matrices,
    const float *projection_multiplier, int D, int H, int W, int hardware_interp, int mode,
    int z_begin, int z_end, float *volume, void *workspace, size_t workspace_bytes,
    pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(sinogram && volume_origin && volume_spacing && projection_matrices &&
                       projection_multiplier && volume && workspace,
                   "cone_backprojection3d_views: null pointer");
  PYRONN_CHECK_ARG(dtype_ok(sinogram_dtype), "cone_backprojection3d_views: bad sinogram dtype %d",
                   sinogram_dtype);
  PYRONN_CHECK_ARG(mode >= 0 && mode <= 7, "cone_backprojection3d_views: bad mode %d", mode);
  const int rc = bp_check_shape("cone_backprojection3d_views", batch, D, H, W, n_views, det_h, det_w);
  if (rc != PYRONN_OK) return rc;
  PYRONN_CHECK_ARG(z_begin >= 0 && z_begin < z_end && z_end <= D,
                   "cone_backprojection3d_views: bad z range [%d, %d) for depth %d", z_begin, z_end, D);
  const size_t need =
      pyronn_cone_backprojection3d_workspace_size(batch, n_views, det_h, det_w, sinogram_dtype);
  if (workspace_bytes < need) {
    set_error("cone_backprojection3d_views: workspace %zu bytes < required %zu", workspace_bytes,
              need);
    return PYRONN_ERR_WORKSPACE;
  }
  PYRONN_CHECK_ARG(((uintptr_t)workspace & 15) == 0, "cone_backprojection3d_views: workspace must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const bool hw = hardware_interp != 0;
  const int images = batch * n_views;
  const int flags = ((mode & 1) ? kBpContinue : 0) | ((mode & 2) ? kBpScale : 0);
  if (sinogram_dtype == PYRONN_F32) {
    if (!(mode & 4)) launch_det_cells<float>(sinogram, workspace, images, det_h, det_w, s);
    launch_bp<float>(hw, workspace, volume, projection_matrices, projection_multiplier,
                     volume_spacing, volume_origin, batch, D, H, W, n_views, det_h, det_w, s, flags, z_begin,
                     z_end);
  } else if (sinogram_dtype == PYRONN_BF16) {
    if (!(mode & 4)) launch_det_cells<bf16s>(sinogram, workspace, images, det_h, det_w, s);
    launch_bp<bf16s>(hw, workspace, volume, projection_matrices, projection_multiplier,
                     volume_spacing, volume_origin, batch, D, H, W, n_views, det_h, det_w, s, flags, z_begin,
                     z_end);
  } else {
    if (!(mode & 4)) launch_det_cells<_Float16>(sinogram, workspace, images, det_h, det_w, s);
    launch_bp<_Float16>(hw, workspace, volume, projection_matrices, projection_multiplier,
                        volume_spacing, volume_origin, batch, D, H, W, n_views, det_h, det_w, s,
                        flags, z_begin, z_end);
  }
  PYRONN_CHECK_LAUNCH("cone_bp_kernel");
  return PYRONN_OK;
}

// =====================================================================================
// Batched launches over problems of different shapes
// =====================================================================================
namespace pyronn {

static int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 256;
  return cus;
}

static int check_problems(const char *name, const pyronn_cone_problem *pr, int count, int dtype,
                          bool forward) {
  PYRONN_CHECK_ARG(pr != nullptr, "%s: null problem list", name);
  PYRONN_CHECK_ARG(count >= 1 && count <= kMaxBatch, "%s: count %d not in [1, %d]", name, count,
                   kMaxBatch);
  PYRONN_CHECK_ARG(dtype_ok(dtype), "%s: bad dtype %d", name, dtype);
  for (int k = 0; k < count; ++k) {
    const pyronn_cone_problem &p = pr[k];
    PYRONN_CHECK_ARG(p.input && p.output && p.volume_origin && p.volume_spacing &&
                         p.projection_matrices && (forward ? p.step_size != nullptr
                                                           : p.projection_multiplier != nullptr),
                     "%s: null pointer in problem %d", name, k);
    PYRONN_CHECK_ARG(p.D > 0 && p.H > 0 && p.W > 0 && p.n_views > 0 && p.det_h > 0 && p.det_w > 0,
                     "%s: bad shape in problem %d", name, k);
    if (forward) {
      if (!fwd_volume_ok(p.D, p.H, p.W)) {
        set_error("%s: problem %d: volume %dx%dx%d has a side of 2^24 - 3 voxels or more", name, k, p.D, p.H,
                  p.W);
        return PYRONN_ERR_TOO_LARGE;
      }
    } else {
      const int rc = bp_check_shape(name, 1, p.D, p.H, p.W, p.n_views, p.det_h, p.det_w);
      if (rc != PYRONN_OK) return rc;
    }
  }
  return PYRONN_OK;
}

static size_t fwd_problem_bytes(const pyronn_cone_problem &p, int dtype) {
  const size_t e = dtype_bytes(dtype);
  return fwd_geom_bytes(1, p.n_views) + align_up(cell_elems_dt(dtype, p.D, p.H, p.W) * e, 256);
}

constexpr size_t kCounterBytes = 256;  // per-XCD work counters at the start of a batched workspace

// g.x > 0: a persistent grid of g.x workgroups claiming items; g.x == 0: one workgroup per
// item, in launches of at most kMaxGridX items (a multiple of 8).
template <typename TS, bool HW, typename IT>
static void launch_fwd_persistent(const FwdBatch &fb, int tile, dim3 g, unsigned long long *counters,
                                  hipStream_t s) {
  if (g.x == 0) {
    const long long per = kMaxGridX / kXcds * kXcds;
    for (long long i0 = 0; i0 < fb.total; i0 += per) {
      const dim3 gi((unsigned)(fb.total - i0 < per ? fb.total - i0 : per));
      switch (tile) {
        case 1: cone_fwd_items_kernel<TS, HW, IT, FwdTile<8, 32, 2>><<<gi, dim3(256), 0, s>>>(fb, i0); break;
        case 2: cone_fwd_items_kernel<TS, HW, IT, FwdTile<32, 8, 8>><<<gi, dim3(256), 0, s>>>(fb, i0); break;
        case 3: cone_fwd_items_kernel<TS, HW, IT, FwdTile<16, 16, 8>><<<gi, dim3(256), 0, s>>>(fb, i0); break;
        default: cone_fwd_items_kernel<TS, HW, IT, FwdTile16><<<gi, dim3(256), 0, s>>>(fb, i0); break;
      }
    }
    return;
  }
  switch (tile) {
    case 1: cone_fwd_persistent_kernel<TS, HW, IT, FwdTile<8, 32, 2>><<<g, dim3(256), 0, s>>>(fb, counters); break;
    case 2: cone_fwd_persistent_kernel<TS, HW, IT, FwdTile<32, 8, 8>><<<g, dim3(256), 0, s>>>(fb, counters); break;
    case 3: cone_fwd_persistent_kernel<TS, HW, IT, FwdTile<16, 16, 8>><<<g, dim3(256), 0, s>>>(fb, counters); break;
    default: cone_fwd_persistent_kernel<TS, HW, IT, FwdTile16><<<g, dim3(256), 0, s>>>(fb, counters); break;
  }
}

template <typename TS>
static int fwd_batched(const pyronn_cone_problem *pr, int count, bool hw, int bpc, int tile, char *ws,
                       hipStream_t s) {
  unsigned long long *counters = (unsigned long long *)ws;
  ws += kCounterBytes;
  FwdBatch fb{};
  long long begin = 0;
  bool small = true, bytes = sizeof(TS) == 2;
  const int tc = fwd_tile_cols(tile), tr = fwd_tile_rows(tile);
  for (int k = 0; k < count; ++k) {
    const pyronn_cone_problem &p = pr[k];
    float *inv_ar = (float *)ws;
    float *src = inv_ar + (size_t)p.n_views * 9;
    char *cells = ws + fwd_geom_bytes(1, p.n_views);
    ws += fwd_problem_bytes(p, sizeof(TS) == 4 ? PYRONN_F32 : PYRONN_BF16);
    cone_setup_kernel<<<dim3((p.n_views + 127) / 128), dim3(128), 0, s>>>(
        p.projection_matrices, p.volume_origin, p.volume_spacing, 1, p.n_views, inv_ar, src);
    launch_cells<TS>(p.input, cells, 1, p.D, p.H, p.W, s);
    small = small && fwd_small_index(p.D, p.H, p.W);
    bytes = bytes && fwd_byte_offsets(p.D, p.H, p.W);
    fb.p[k] = FwdProblem{cells, p.output, inv_ar, src, p.step_size, p.volume_spacing,
                         p.W, p.H, p.D, p.n_views, p.det_h, p.det_w,
                         (p.det_w + tc - 1) / tc, (p.det_h + tr - 1) / tr, begin};
    begin += fwd_blocks(p.n_views, p.det_h, p.det_w, tile);
  }
  fb.count = count;
  fb.total = begin;
  // bpc <= 0: one workgroup per item (grid 0 below); else a persistent grid
  long long grid = 0;
  if (bpc > 0) {
    grid = (long long)device_cus() * bpc;
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    if (grid > begin) grid = begin;
    reset_counters_kernel<<<1, 64, 0, s>>>(counters);
  }
  const dim3 g((unsigned)grid);
  if constexpr (sizeof(TS) == 2) {
    if (bytes) {
      if (hw) launch_fwd_persistent<TS, true, unsigned>(fb, tile, g, counters, s);
      else launch_fwd_persistent<TS, false, unsigned>(fb, tile, g, counters, s);
      return PYRONN_OK;
    }
  }
  if (small) {
    if (hw) launch_fwd_persistent<TS, true, int>(fb, tile, g, counters, s);
    else launch_fwd_persistent<TS, false, int>(fb, tile, g, counters, s);
  } else {
    if (hw) launch_fwd_persistent<TS, true, long long>(fb, tile, g, counters, s);
    else launch_fwd_persistent<TS, false, long long>(fb, tile, g, counters, s);
  }
  return PYRONN_OK;
}

template <typename TS, int Z>
static int bp_batched(const pyronn_cone_problem *pr, int count, bool hw, int bpc, char *ws,
                      hipStream_t s) {
  unsigned long long *counters = (unsigned long long *)ws;
  ws += kCounterBytes;
  BpBatch bb{};
  long long begin = 0;
  for (int k = 0; k < count; ++k) {
    const pyronn_cone_problem &p = pr[k];
    char *cells = ws;
    ws += align_up(det_cells(p.n_views, p.det_h, p.det_w) * 4 * sizeof(TS), 256);
    launch_det_cells<TS>(p.input, cells, p.n_views, p.det_h, p.det_w, s);
    const int zblocks = (p.D + Z - 1) / Z;
    const int tx = (p.W + 15) / 16, ty = (p.H + 15) / 16;
    bb.p[k] = BpProblem{cells, p.output, p.projection_matrices, p.projection_multiplier,
                        p.volume_spacing, p.volume_origin, p.W, p.H, p.D, p.n_views, p.det_h,
                        p.det_w, tx, ty, zblocks, begin, kBpScale, 0, p.D};
    begin += (long long)bp_items_per_z(tx, ty) * zblocks;  // a multiple of 8
  }
  bb.count = count;
  bb.total = begin;
  bool fidx = true;
  for (int k = 0; k < count; ++k) fidx = fidx && det_fidx(pr[k].det_h, pr[k].det_w);
  if (bpc <= 0 && begin <= kMaxGridX) {  // one workgroup per item (the default; else the persistent grid)
    const dim3 g((unsigned)begin);
    if (fidx) {
      if (hw) cone_bp_items_kernel<TS, true, Z, true><<<g, dim3(256), 0, s>>>(bb);
      else cone_bp_items_kernel<TS, false, Z, true><<<g, dim3(256), 0, s>>>(bb);
    } else {
      if (hw) cone_bp_items_kernel<TS, true, Z, false><<<g, dim3(256), 0, s>>>(bb);
      else cone_bp_items_kernel<TS, false, Z, false><<<g, dim3(256), 0, s>>>(bb);
    }
    return PYRONN_OK;
  }
  long long grid = (long long)device_cus() * (bpc > 0 ? bpc : 8);
  grid = (grid + kXcds - 1) / kXcds * kXcds;
  reset_counters_kernel<<<1, 64, 0, s>>>(counters);
  const dim3 g((unsigned)grid);
  if (fidx) {
    if (hw) cone_bp_persistent_kernel<TS, true, Z, true><<<g, dim3(256), 0, s>>>(bb, counters);
    else cone_bp_persistent_kernel<TS, false, Z, true><<<g, dim3(256), 0, s>>>(bb, counters);
  } else {
    if (hw) cone_bp_persistent_kernel<TS, true, Z, false><<<g, dim3(256), 0, s>>>(bb, counters);
    else cone_bp_persistent_kernel<TS, false, Z, false><<<g, dim3(256), 0, s>>>(bb, counters);
  }
  return PYRONN_OK;
}

}  // namespace pyronn

extern "C" size_t pyronn_cone_projection3d_batched_workspace_size(const pyronn_cone_problem *problems,
                                                                  int count, int dtype) {
  if (!problems || count < 1 || count > kMaxBatch || !dtype_ok(dtype)) return 0;
  size_t total = kCounterBytes;
  for (int k = 0; k < count; ++k) total += fwd_problem_bytes(problems[k], dtype);
  return total;
}

extern "C" int pyronn_cone_projection3d_batched(const pyronn_cone_problem *problems, int count,
                                                int dtype, int hardware_interp, int blocks_per_cu,
                                                int forward_tile, void *workspace, size_t workspace_bytes,
                                                pyronn_stream_t stream) {
  clear_error();
  int rc = check_problems("cone_projection3d_batched", problems, count, dtype, true);
  if (rc != PYRONN_OK) return rc;
  PYRONN_CHECK_ARG(forward_tile >= 0 && forward_tile < kFwdTiles,
                   "cone_projection3d_batched: forward_tile %d not in [0, %d)", forward_tile, kFwdTiles);
  PYRONN_CHECK_ARG(workspace != nullptr, "cone_projection3d_batched: null workspace");
  const size_t need = pyronn_cone_projection3d_batched_workspace_size(problems, count, dtype);
  if (workspace_bytes < need) {
    set_error("cone_projection3d_batched: workspace %zu bytes < required %zu", workspace_bytes, need);
    return PYRONN_ERR_WORKSPACE;
  }
  PYRONN_CHECK_ARG(((uintptr_t)workspace & 15) == 0, "cone_projection3d_batched: workspace must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const bool hw = hardware_interp != 0;
  char *ws = (char *)workspace;
  if (dtype == PYRONN_F32) rc = fwd_batched<float>(problems, count, hw, blocks_per_cu, forward_tile, ws, s);
  else if (dtype == PYRONN_BF16) rc = fwd_batched<bf16s>(problems, count, hw, blocks_per_cu, forward_tile, ws, s);
  else rc = fwd_batched<_Float16>(problems, count, hw, blocks_per_cu, forward_tile, ws, s);
  PYRONN_CHECK_LAUNCH("cone_fwd_persistent_kernel");
  return rc;
}

extern "C" size_t pyronn_cone_backprojection3d_batched_workspace_size(
    const pyronn_cone_problem *problems, int count, int dtype) {
  if (!problems || count < 1 || count > kMaxBatch || !dtype_ok(dtype)) return 0;
  size_t total = kCounterBytes;
  for (int k = 0; k < count; ++k)
    total += align_up(det_cells(problems[k].n_views, problems[k].det_h, problems[k].det_w) * 4 *
                          dtype_bytes(dtype),
                      256);
  return total;
}

extern "C" int pyronn_cone_backprojection3d_batched(const pyronn_cone_problem *problems, int count,
                                                    int dtype, int hardware_interp, int z_per_lane,
                                                    int blocks_per_cu, void *workspace,
                                                    size_t workspace_bytes, pyronn_stream_t stream) {
  clear_error();
  int rc = check_problems("cone_backprojection3d_batched", problems, count, dtype, false);
  if (rc != PYRONN_OK) return rc;
  PYRONN_CHECK_ARG(workspace != nullptr, "cone_backprojection3d_batched: null workspace");
  PYRONN_CHECK_ARG(z_per_lane <= 0 || z_per_lane == 4 || z_per_lane == 8,
                   "cone_backprojection3d_batched: z_per_lane must be 4 or 8");
  const size_t need = pyronn_cone_backprojection3d_batched_workspace_size(problems, count, dtype);
  if (workspace_bytes < need) {
    set_error("cone_backprojection3d_batched: workspace %zu bytes < required %zu", workspace_bytes,
              need);
    return PYRONN_ERR_WORKSPACE;
  }
  PYRONN_CHECK_ARG(((uintptr_t)workspace & 15) == 0, "cone_backprojection3d_batched: workspace must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const bool hw = hardware_interp != 0;
  const int z = z_per_lane > 0 ? z_per_lane : (dtype == PYRONN_F16 ? 4 : 8);
  char *ws = (char *)workspace;
  if (dtype == PYRONN_F32)
    rc = z == 8 ? bp_batched<float, 8>(problems, count, hw, blocks_per_cu, ws, s)
                : bp_batched<float, 4>(problems, count, hw, blocks_per_cu, ws, s);
  else if (dtype == PYRONN_BF16)
    rc = z == 8 ? bp_batched<bf16s, 8>(problems, count, hw, blocks_per_cu, ws, s)
                : bp_batched<bf16s, 4>(problems, count, hw, blocks_per_cu, ws, s);
  else
    rc = z == 8 ? bp_batched<_Float16, 8>(problems, count, hw, blocks_per_cu, ws, s)
                : bp_batched<_Float16, 4>(problems, count, hw, blocks_per_cu, ws, s);
  PYRONN_CHECK_LAUNCH("cone_bp_persistent_kernel");
  return rc;
}

