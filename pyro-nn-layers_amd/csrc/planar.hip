// Parallel- and fan-beam 2D projector / backprojector for gfx950.
//
// The reference runs all four 2D kernels on the CUDA texture unit (linear
// filter, border addressing, unnormalised coordinates).  gfx950 has no image
// sampler, so the texture fetch is done in software: the coordinate c handed
// to the sampler is formed exactly as the reference forms it
// (physical_to_index(p) + 0.5f), then b = c - 0.5f, i = floor(b), weights
// b - i (exact, not 8-bit), neighbours outside the array read 0.
//
// The 2D problems are small (a 512^2 image is 1 MiB), so the samplers test
// bounds per neighbour instead of reading a zero-apron copy.  With f32 storage the two
// x neighbours of a row are adjacent: when both lie inside the row, one 8-byte load reads
// them (dword aligned; the target's vector loads do not need natural alignment), which
// halves the loads of a sample without changing any value.
#include "common.h"

namespace pyronn {

typedef float f32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));

// img[i], img[i + 1] as one load
__device__ __forceinline__ f32x2_a4 load_pair(const float *__restrict__ img, size_t i) {
  return *reinterpret_cast<const f32x2_a4 *>(img + i);
}

constexpr float kPi = 3.14159265359f;  // par_backprojector_2D_CudaKernel.cu.cc:32

// helper_grid.h:23-31
__device__ __forceinline__ float i2p(float i, float o, float s) { return fmaf(i, s, o); }
__device__ __forceinline__ float p2i(float p, float o, float s) { return (p - o) / s; }

// 1 / s when s is a positive power of two well inside the normal range, else 0.  For
// such s, x / s and x * (1 / s) are the correctly rounded value of the same real number,
// so they are equal for every x; the kernels then multiply instead of dividing.
__device__ __forceinline__ float pow2_reciprocal(float s) {
  const unsigned b = __float_as_uint(s);
  const unsigned e = (b >> 23) & 0xffu;
  return ((b & 0x807fffffu) == 0u && e > 1u && e < 253u) ? __uint_as_float((254u - e) << 23) : 0.0f;
}
// physical_to_index with the division replaced by the exact product when POW2
template <bool POW2>
__device__ __forceinline__ float p2i_k(float p, float o, float s, float inv) {
  return POW2 ? (p - o) * inv : (p - o) / s;
}

// The four neighbours a lane read last (consecutive march samples usually share them).
struct Tex2dRegs {
  int x0 = 0x7fffffff, y0 = 0;
  float p00 = 0.f, p01 = 0.f, p10 = 0.f, p11 = 0.f;
};

// tex2d with the neighbours kept in registers across calls: loads only when (x0, y0)
// changes; the same values and the same blend as tex2d.
template <typename TS>
__device__ __forceinline__ float tex2d_cached(const TS *__restrict__ img, int W, int H, float cx,
                                              float cy, Tex2dRegs &r) {
  float bx = cx - 0.5f, by = cy - 0.5f;
  if (!(bx > -2.0f && bx < (float)(W + 1) && by > -2.0f && by < (float)(H + 1))) return 0.0f;
  const float fx = floorf(bx), fy = floorf(by);
  const int x0 = (int)fx, y0 = (int)fy;
  if (x0 != r.x0 || y0 != r.y0) {
    r.x0 = x0;
    r.y0 = y0;
    const int x1 = x0 + 1, y1 = y0 + 1;
    const bool vx0 = x0 >= 0 && x0 < W, vx1 = x1 >= 0 && x1 < W;
    const bool vy0 = y0 >= 0 && y0 < H, vy1 = y1 >= 0 && y1 < H;
    if (sizeof(TS) == 4 && vx0 && vx1) {
      const float *f = reinterpret_cast<const float *>(img);
      const f32x2_a4 a = vy0 ? load_pair(f, (size_t)y0 * W + x0) : f32x2_a4{0.0f, 0.0f};
      const f32x2_a4 b = vy1 ? load_pair(f, (size_t)y1 * W + x0) : f32x2_a4{0.0f, 0.0f};
      r.p00 = a.x;
      r.p01 = a.y;
      r.p10 = b.x;
      r.p11 = b.y;
    } else {
      r.p00 = (vx0 && vy0) ? to_f32(img[y0 * W + x0]) : 0.0f;
      r.p01 = (vx1 && vy0) ? to_f32(img[y0 * W + x1]) : 0.0f;
      r.p10 = (vx0 && vy1) ? to_f32(img[y1 * W + x0]) : 0.0f;
      r.p11 = (vx1 && vy1) ? to_f32(img[y1 * W + x1]) : 0.0f;
    }
  }
  const float xd = bx - fx, yd = by - fy;
  const float p0 = lerp_ref(r.p00, r.p01, xd);
  const float p1 = lerp_ref(r.p10, r.p11, xd);
  return lerp_ref(p0, p1, yd);
}

template <typename TS>
__device__ __forceinline__ float tex1d(const TS *__restrict__ row, int W, float cx) {
  const float bx = cx - 0.5f;
  if (!(bx > -2.0f && bx < (float)(W + 1))) return 0.0f;
  const float fx = floorf(bx);
  const int x0 = (int)fx, x1 = x0 + 1;
  float p0, p1;
  if (sizeof(TS) == 4 && x0 >= 0 && x1 < W) {
    const f32x2_a4 a = load_pair(reinterpret_cast<const float *>(row), (size_t)x0);
    p0 = a.x;
    p1 = a.y;
  } else {
    p0 = (x0 >= 0 && x0 < W) ? to_f32(row[x0]) : 0.0f;
    p1 = (x1 >= 0 && x1 < W) ? to_f32(row[x1]) : 0.0f;
  }
  return lerp_ref(p0, p1, bx - fx);
}

// Ray parameter interval inside the image box, whose edges sit 0.5 mm before each index
// edge (par_projector_2D_CudaKernel.cu.cc:33-57); an axis-parallel ray skips that slab.
__device__ __forceinline__ void clip2d(int W, int H, float ox, float oy, float sx, float sy,
                                       float s0, float s1, float r0, float r1, float &amin,
                                       float &amax) {
  amin = 0.0f;
  amax = __builtin_inff();
  if (0.0f != r0) {
    const float e0 = i2p(0.0f, ox, sx) - 0.5f, e1 = i2p((float)W, ox, sx) - 0.5f;
    const float reci = 1.0f / r0;
    const float a0 = (e0 - s0) * reci, a1 = (e1 - s0) * reci;
    amin = fminf(a0, a1);
    amax = fmaxf(a0, a1);
  }
  if (0.0f != r1) {
    const float e0 = i2p(0.0f, oy, sy) - 0.5f, e1 = i2p((float)H, oy, sy) - 0.5f;
    const float reci = 1.0f / r1;
    const float a0 = (e0 - s1) * reci, a1 = (e1 - s1) * reci;
    amin = fmaxf(amin, fminf(a0, a1));
    amax = fminf(amax, fmaxf(a0, a1));
  }
}

// kernel_project2D, par_projector_2D_CudaKernel.cu.cc:28-102 (same code in
// fan_projector_2D_CudaKernel.cu.cc:27-102): march in mm, box edges with the
// -0.5 mm offset, first sample exactly at the entry point, tail for a positive sum.
template <typename TS, bool POW2>
__device__ float march2d(const TS *__restrict__ img, int W, int H, float ox, float oy, float sx,
                         float sy, float ix, float iy, float s0, float s1, float r0, float r1,
                         float step) {
  if (!(step > 0.0f)) return __builtin_nanf("");
  float amin, amax;
  clip2d(W, H, ox, oy, sx, sy, s0, s1, r0, r1, amin, amax);
  float pix = 0.0f, px = 0.0f, py = 0.0f;
  Tex2dRegs regs;
  const int cap = march_cap(amin, amax, step);
#define PYRONN_TEX2D(X, Y) \
  tex2d_cached(img, W, H, p2i_k<POW2>((X), ox, sx, ix) + 0.5f, p2i_k<POW2>((Y), oy, sy, iy) + 0.5f, regs)
  if (amin < amax) {
    px = fmaf(amin, r0, s0);
    py = fmaf(amin, r1, s1);
    pix = fmaf(0.5f, PYRONN_TEX2D(px, py), pix);
    amin += step;
  }
  for (int it = 0; amin < amax && it < cap; ++it) {
    px = fmaf(amin, r0, s0);
    py = fmaf(amin, r1, s1);
    pix += PYRONN_TEX2D(px, py);
    amin += step;
  }
  pix *= step;
  if (pix > 0.0f) {
    const float flast = PYRONN_TEX2D(px, py);
    pix = fmaf(-(0.5f * step), flast, pix);
    amin -= step;
    const float last = amax - amin;
    pix = fmaf(0.5f * last, flast, pix);
    px = fmaf(amax, r0, s0);
    py = fmaf(amax, r1, s1);
    pix = fmaf(0.5f * last, PYRONN_TEX2D(px, py), pix);
  }
#undef PYRONN_TEX2D
  return pix;
}

// Value of lane J of this lane's quad (DPP quad_perm broadcast; all four lanes active).
template <int J>
__device__ __forceinline__ float quad_bcast(float f) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(f), J * 0x55, 0xf, 0xf, false));
}

// The march of march2d with the samples of one ray spread over the G = 4 lanes of a quad:
// every lane runs the same alpha recurrence (alpha += step, the same float adds), lane q
// samples positions q, q + G, ... of each group of G, and every lane adds the G values in
// sample order from quad broadcasts, so all G lanes hold the same running sum, bit for bit
// the one-lane sum.  The entrance sample and the tail are computed by all lanes alike.
// Four times the lanes per ray: for launches with too few rays to fill the GPU, where the
// one-lane march is bound by its own load latency (measured on 256^2 / 180 views x 363
// pixels: 0.67 -> 0.35 ms; with 261K rays, 2 or 4 lanes per ray were slower than one).
template <typename TS, bool POW2, int G>
__device__ float march2d_group(const TS *__restrict__ img, int W, int H, float ox, float oy,
                               float sx, float sy, float ix, float iy, float s0, float s1, float r0,
                               float r1, float step, int q) {
  if (!(step > 0.0f)) return __builtin_nanf("");
  float amin, amax;
  clip2d(W, H, ox, oy, sx, sy, s0, s1, r0, r1, amin, amax);
  float pix = 0.0f, alast = 0.0f;
  bool any = false;  // a sample was taken: (px, py) of the one-lane march is at alast
  Tex2dRegs regs;
  const int cap = march_cap(amin, amax, step);
#define PYRONN_TEX2D(X, Y) \
  tex2d_cached(img, W, H, p2i_k<POW2>((X), ox, sx, ix) + 0.5f, p2i_k<POW2>((Y), oy, sy, iy) + 0.5f, regs)
  if (amin < amax) {
    pix = fmaf(0.5f, PYRONN_TEX2D(fmaf(amin, r0, s0), fmaf(amin, r1, s1)), pix);
    alast = amin;
    any = true;
    amin += step;
  }
  for (int it = 0;;) {
    float a[G + 1];
    a[0] = amin;
#pragma unroll
    for (int j = 1; j <= G; ++j) a[j] = a[j - 1] + step;
    // sample j of the group exists iff the one-lane loop would take it: a prefix, as
    // alpha only grows
    bool ok[G];
#pragma unroll
    for (int j = 0; j < G; ++j) ok[j] = (j == 0 || ok[j - 1]) && a[j] < amax && it + j < cap;
    float my_a = a[0];
    bool my_ok = ok[0];
#pragma unroll
    for (int j = 1; j < G; ++j) {
      my_a = q == j ? a[j] : my_a;
      my_ok = q == j ? ok[j] : my_ok;
    }
    float f = 0.0f;
    if (my_ok) f = PYRONN_TEX2D(fmaf(my_a, r0, s0), fmaf(my_a, r1, s1));
    static_assert(G == 4, "quad broadcasts");
    const float fg[G] = {quad_bcast<0>(f), quad_bcast<1>(f), quad_bcast<2>(f), quad_bcast<3>(f)};
    int c = 0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (ok[j]) pix += fg[j];
      c += ok[j] ? 1 : 0;
    }
    if (c > 0) {
      alast = a[c - 1];
      any = true;
    }
    amin = a[c];
    it += c;
    if (c < G) break;
  }
  pix *= step;
  if (pix > 0.0f) {
    float px = 0.0f, py = 0.0f;
    if (any) {
      px = fmaf(alast, r0, s0);
      py = fmaf(alast, r1, s1);
    }
    const float flast = PYRONN_TEX2D(px, py);
    pix = fmaf(-(0.5f * step), flast, pix);
    amin -= step;
    const float last = amax - amin;
    pix = fmaf(0.5f * last, flast, pix);
    px = fmaf(amax, r0, s0);
    py = fmaf(amax, r1, s1);
    pix = fmaf(0.5f * last, PYRONN_TEX2D(px, py), pix);
  }
#undef PYRONN_TEX2D
  return pix;
}

// One lane per (detector pixel, view, batch item).
// Parallel: project_2Dpar_beam_kernel, par_projector_2D_CudaKernel.cu.cc:104-147.
// Fan:      project_2Dfan_beam_kernel, fan_projector_2D_CudaKernel.cu.cc:104-149.
// G lanes per pixel (1, or 4 with march2d_group for launches with few rays).
template <typename TS, bool FAN, int G>
__global__ void __launch_bounds__(256)
proj2d_kernel(const TS *__restrict__ vol, float *__restrict__ sino, const float *__restrict__ rays,
              const float *__restrict__ vorigin, const float *__restrict__ dorigin,
              const float *__restrict__ vspacing, const float *__restrict__ dspacing,
              const float *__restrict__ sid_p, const float *__restrict__ sdd_p, int H, int W, int n,
              int det, float step_factor) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = t / G, q = t % G;
  const int view = blockIdx.y, b = blockIdx.z;
  if (d >= det) return;  // the G lanes of a pixel share d
  const float sx = vspacing[b * 2 + 1], sy = vspacing[b * 2 + 0];
  const float ox = vorigin[b * 2 + 1], oy = vorigin[b * 2 + 0];
  const size_t vb = (size_t)b * n + view;
  const float c0 = rays[vb * 2 + 0], c1 = rays[vb * 2 + 1];
  const float u0 = -c1, u1 = c0;
  const float u = i2p((float)d, dorigin[b], dspacing[b]);
  float s0, s1, r0, r1;
  if (!FAN) {
    const float a = (((float)W * sx) * (float)W) * sx;
    const float sid = sqrtf(fmaf(((float)H * sy) * (float)H, sy, a)) * 1.2f;
    s0 = fmaf(u0, u, c0 * (-sid));
    s1 = fmaf(u1, u, c1 * (-sid));
    r0 = c0;
    r1 = c1;
  } else {
    const float sid = sid_p[b], sdd = sdd_p[b];
    s0 = c0 * (-sid);
    s1 = c1 * (-sid);
    const float dx = fmaf(u0, u, fmaf(c0, sdd, s0));
    const float dy = fmaf(u1, u, fmaf(c1, sdd, s1));
    const float ex = dx - s0, ey = dy - s1;
    const float nrm = 1.0f / sqrtf(fmaf(ey, ey, ex * ex));
    r0 = ex * nrm;
    r1 = ey * nrm;
  }
  const float step = step_factor * fminf(sx, sy);
  const TS *img = vol + (size_t)b * H * W;
  // spacings are per batch item, so the choice is uniform over the block
  const float ix = pow2_reciprocal(sx), iy = pow2_reciprocal(sy);
  float pix;
  if constexpr (G > 1)
    pix = (ix != 0.0f && iy != 0.0f)
              ? march2d_group<TS, true, G>(img, W, H, ox, oy, sx, sy, ix, iy, s0, s1, r0, r1, step, q)
              : march2d_group<TS, false, G>(img, W, H, ox, oy, sx, sy, ix, iy, s0, s1, r0, r1, step, q);
  else
    pix = (ix != 0.0f && iy != 0.0f)
              ? march2d<TS, true>(img, W, H, ox, oy, sx, sy, ix, iy, s0, s1, r0, r1, step)
              : march2d<TS, false>(img, W, H, ox, oy, sx, sy, ix, iy, s0, s1, r0, r1, step);
  const float a = r0 * sx, c = r1 * sy;
  if (q == 0) sino[vb * det + d] = pix * sqrtf(fmaf(c, c, a * a));
}

// 2D line intersection, helper_geometry_gpu.h:28-44 (NaN when |den| < 1e-6).
__device__ __forceinline__ void isect2(float p1x, float p1y, float p2x, float p2y, float p3x,
                                       float p3y, float p4x, float p4y, float &ox, float &oy) {
  const float a = p1x - p2x, b = p3y - p4y, c = p1y - p2y, d = p3x - p4x;
  const float den = fmaf(-c, d, a * b);
  if (den < 0.000001f && den > -0.000001f) {
    ox = oy = __builtin_nanf("");
    return;
  }
  const float e = fmaf(-p1y, p2x, p1x * p2y);
  const float g = fmaf(-p3y, p4x, p3x * p4y);
  ox = fmaf(-a, g, e * d) / den;
  oy = fmaf(-c, g, e * b) / den;
}

// Sum over all views for the pixel at physical (pcx, pcy).
template <typename TS, bool FAN, bool POW2>
__device__ __forceinline__ float back2d_sum(const TS *__restrict__ sb, const float *__restrict__ rb, int n,
                                            int det, float pcx, float pcy, float do_, float ds,
                                            float ids, float sid, float sdd) {
  float val = 0.0f;
  if (!FAN) {
    for (int v = 0; v < n; ++v) {
      const float c0 = rb[v * 2 + 0], c1 = rb[v * 2 + 1];
      const float s = fmaf(pcy, c0, pcx * (-c1));
      const float si = p2i_k<POW2>(s, do_, ds, ids);
      val += tex1d(sb + (size_t)v * det, det, si + 0.5f);
    }
  } else {
    for (int v = 0; v < n; ++v) {
      const float c0 = rb[v * 2 + 0], c1 = rb[v * 2 + 1];
      const float d0 = -c1, d1 = c0;
      const float srx = c0 * (-sid), sry = c1 * (-sid);
      const float cpx = fmaf(c0, sdd, srx), cpy = fmaf(c1, sdd, sry);
      float ix, iy;
      isect2(pcx, pcy, srx, sry, cpx, cpy, cpx + d0, cpy + d1, ix, iy);
      const float ex = pcx - srx, ey = pcy - sry;
      // 1 / |x - source| as the reference's division (rcp_rn equals 1.0f / w for every w)
      const float dw = rcp_rn(sqrtf(fmaf(ey, ey, ex * ex)));
      const float s = fmaf(iy, d1, ix * d0);
      const float si = p2i_k<POW2>(s, do_, ds, ids);
      const float f = tex1d(sb + (size_t)v * det, det, si + 0.5f);
      val = fmaf(f * dw, dw, val);
    }
  }
  return val;
}

// One lane per pixel; loop over all views.
// Parallel: backproject_2Dpar_beam_kernel, par_backprojector_2D_CudaKernel.cu.cc:28-61.
// Fan:      backproject_2Dfan_beam_kernel, fan_backprojector_2D_CudaKernel.cu.cc:29-67.
template <typename TS, bool FAN>
__global__ void __launch_bounds__(256)
back2d_kernel(const TS *__restrict__ sino, float *__restrict__ vol, const float *__restrict__ rays,
              const float *__restrict__ vorigin, const float *__restrict__ dorigin,
              const float *__restrict__ vspacing, const float *__restrict__ dspacing,
              const float *__restrict__ sid_p, const float *__restrict__ sdd_p, int H, int W, int n,
              int det) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  const int b = blockIdx.z;
  if (x >= W || y >= H) return;
  const float sx = vspacing[b * 2 + 1], sy = vspacing[b * 2 + 0];
  const float ox = vorigin[b * 2 + 1], oy = vorigin[b * 2 + 0];
  const float do_ = dorigin[b], ds = dspacing[b];
  const float pcx = i2p((float)x, ox, sx), pcy = i2p((float)y, oy, sy);
  const float *rb = rays + (size_t)b * n * 2;
  const TS *sb = sino + (size_t)b * n * det;
  // the detector spacing is per batch item, so the choice is uniform over the block
  const float ids = pow2_reciprocal(ds);
  float sid = 0.0f, sdd = 0.0f;
  if (FAN) {  // the parallel op passes no source distances
    sid = sid_p[b];
    sdd = sdd_p[b];
  }
  const float val = ids != 0.0f ? back2d_sum<TS, FAN, true>(sb, rb, n, det, pcx, pcy, do_, ds, ids, sid, sdd)
                                : back2d_sum<TS, FAN, false>(sb, rb, n, det, pcx, pcy, do_, ds, ids, sid, sdd);
  if (!FAN) {
    vol[((size_t)b * H + y) * W + x] = (kPi * val) / (float)n;
  } else {
    vol[((size_t)b * H + y) * W + x] = (((sid * sdd) * kPi) * val) / (float)n;
  }
}

// rays per launch below which a ray gets four lanes
constexpr long long kQuadRayLimit = 131072;

template <bool FAN, int G>
static void launch_proj2d_g(dim3 grid, const void *volume, int dtype, int H, int W, const float *vo,
                            const float *dor, const float *vs, const float *ds, const float *sid,
                            const float *sdd, const float *rays, int n, int det, float step_factor,
                            float *sino, hipStream_t s) {
  if (dtype == PYRONN_F32)
    proj2d_kernel<float, FAN, G><<<grid, dim3(256), 0, s>>>((const float *)volume, sino, rays, vo, dor, vs,
                                                            ds, sid, sdd, H, W, n, det, step_factor);
  else if (dtype == PYRONN_BF16)
    proj2d_kernel<bf16s, FAN, G><<<grid, dim3(256), 0, s>>>((const bf16s *)volume, sino, rays, vo, dor, vs,
                                                            ds, sid, sdd, H, W, n, det, step_factor);
  else
    proj2d_kernel<_Float16, FAN, G><<<grid, dim3(256), 0, s>>>(
        (const _Float16 *)volume, sino, rays, vo, dor, vs, ds, sid, sdd, H, W, n, det, step_factor);
}

template <bool FAN>
static int launch_proj2d(const char *name, const void *volume, int dtype, int batch, int H, int W,
                         const float *vo, const float *dor, const float *vs, const float *ds,
                         const float *sid, const float *sdd, const float *rays_p, int n, int det,
                         float step_factor, float *sino, hipStream_t s) {
  clear_error();
  PYRONN_CHECK_ARG(volume && vo && dor && vs && ds && rays_p && sino, "%s: null pointer", name);
  PYRONN_CHECK_ARG(!FAN || (sid && sdd), "%s: null source distances", name);
  PYRONN_CHECK_ARG(dtype_ok(dtype), "%s: bad dtype %d", name, dtype);
  PYRONN_CHECK_ARG(batch > 0 && batch <= 65535 && H > 0 && W > 0 && n > 0 && n <= 65535 && det > 0,
                   "%s: bad shape B=%d H=%d W=%d n=%d det=%d", name, batch, H, W, n, det);
  PYRONN_CHECK_ARG((long long)H * W < INT32_MAX, "%s: image too large", name);
  PYRONN_CHECK_ARG(step_factor > 0.0f, "%s: step factor must be > 0", name);
  PYRONN_CHECK_ARG(det <= INT32_MAX / 4, "%s: too many detector pixels", name);
  // Few rays (one lane each would leave most SIMDs with a handful of latency-bound waves):
  // four lanes per ray.  Both variants give the same bits.
  const long long nrays = (long long)det * n * batch;
  const int G = nrays < kQuadRayLimit ? 4 : 1;
  const dim3 grid((unsigned)(((long long)G * det + 255) / 256), n, batch);
  if (G == 4)
    launch_proj2d_g<FAN, 4>(grid, volume, dtype, H, W, vo, dor, vs, ds, sid, sdd, rays_p, n, det, step_factor,
                            sino, s);
  else
    launch_proj2d_g<FAN, 1>(grid, volume, dtype, H, W, vo, dor, vs, ds, sid, sdd, rays_p, n, det, step_factor,
                            sino, s);
  PYRONN_CHECK_LAUNCH(name);
  return PYRONN_OK;
}

template <bool FAN>
static int launch_back2d(const char *name, const void *sino, int dtype, int batch, int n, int det,
                         const float *vo, const float *dor, const float *vs, const float *ds,
                         const float *sid, const float *sdd, const float *rays, int H, int W,
                         float *vol, hipStream_t s) {
  clear_error();
  PYRONN_CHECK_ARG(sino && vo && dor && vs && ds && rays && vol, "%s: null pointer", name);
  PYRONN_CHECK_ARG(!FAN || (sid && sdd), "%s: null source distances", name);
  PYRONN_CHECK_ARG(dtype_ok(dtype), "%s: bad dtype %d", name, dtype);
  PYRONN_CHECK_ARG(batch > 0 && batch <= 65535 && H > 0 && W > 0 && n > 0 && det > 0,
                   "%s: bad shape B=%d H=%d W=%d n=%d det=%d", name, batch, H, W, n, det);
  dim3 grid((W + 15) / 16, (H + 15) / 16, batch);
  if (dtype == PYRONN_F32)
    back2d_kernel<float, FAN><<<grid, dim3(256), 0, s>>>((const float *)sino, vol, rays, vo, dor, vs, ds,
                                                         sid, sdd, H, W, n, det);
  else if (dtype == PYRONN_BF16)
    back2d_kernel<bf16s, FAN><<<grid, dim3(256), 0, s>>>((const bf16s *)sino, vol, rays, vo, dor, vs, ds,
                                                         sid, sdd, H, W, n, det);
  else
    back2d_kernel<_Float16, FAN><<<grid, dim3(256), 0, s>>>((const _Float16 *)sino, vol, rays, vo, dor,
                                                            vs, ds, sid, sdd, H, W, n, det);
  PYRONN_CHECK_LAUNCH(name);
  return PYRONN_OK;
}

}  // namespace pyronn

using namespace pyronn;

extern "C" int pyronn_parallel_projection2d(const void *volume, int volume_dtype, int batch, int H,
                                            int W, const float *volume_origin,
                                            const float *detector_origin,
                                            const float *volume_spacing,
                                            const float *detector_spacing,
                                            const float *ray_vectors, int n_views, int det,
                                            float step_factor, float *sinogram,
                                            pyronn_stream_t stream) {
  return launch_proj2d<false>("parallel_projection2d", volume, volume_dtype, batch, H, W,
                              volume_origin, detector_origin, volume_spacing, detector_spacing,
                              nullptr, nullptr, ray_vectors, n_views, det, step_factor, sinogram,
                              (hipStream_t)stream);
}

extern "C" int pyronn_fan_projection2d(const void *volume, int volume_dtype, int batch, int H,
                                       int W, const float *volume_origin,
                                       const float *detector_origin, const float *volume_spacing,
                                       const float *detector_spacing,
                                       const float *source_2_isocenter_distance,
                                       const float *source_2_detector_distance,
                                       const float *central_ray_vectors, int n_views, int det,
                                       float step_factor, float *sinogram,
                                       pyronn_stream_t stream) {
  return launch_proj2d<true>("fan_projection2d", volume, volume_dtype, batch, H, W, volume_origin,
                             detector_origin, volume_spacing, detector_spacing,
                             source_2_isocenter_distance, source_2_detector_distance,
                             central_ray_vectors, n_views, det, step_factor, sinogram,
                             (hipStream_t)stream);
}

extern "C" int pyronn_parallel_backprojection2d(const void *sinogram, int sinogram_dtype,
                                                int batch, int n_views, int det,
                                                const float *volume_origin,
                                                const float *detector_origin,
                                                const float *volume_spacing,
                                                const float *detector_spacing,
                                                const float *ray_vectors, int H, int W,
                                                float *volume, pyronn_stream_t stream) {
  return launch_back2d<false>("parallel_backprojection2d", sinogram, sinogram_dtype, batch, n_views,
                              det, volume_origin, detector_origin, volume_spacing,
                              detector_spacing, nullptr, nullptr, ray_vectors, H, W, volume,
                              (hipStream_t)stream);
}

extern "C" int pyronn_fan_backprojection2d(const void *sinogram, int sinogram_dtype, int batch,
                                           int n_views, int det, const float *volume_origin,
                                           const float *detector_origin,
                                           const float *volume_spacing,
                                           const float *detector_spacing,
                                           const float *source_2_isocenter_distance,
                                           const float *source_2_detector_distance,
                                           const float *central_ray_vectors, int H, int W,
                                           float *volume, pyronn_stream_t stream) {
  return launch_back2d<true>("fan_backprojection2d", sinogram, sinogram_dtype, batch, n_views, det,
                             volume_origin, detector_origin, volume_spacing, detector_spacing,
                             source_2_isocenter_distance, source_2_detector_distance,
                             central_ray_vectors, H, W, volume, (hipStream_t)stream);
}
