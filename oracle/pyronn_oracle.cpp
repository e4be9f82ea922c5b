// CPU oracle for the PYRO-NN-Layers projector / backprojector path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the arithmetic of the
// reference CUDA kernels for the host so that the HIP library can be checked
// against it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it; the product library never links or calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  Arithmetic conventions shared by every
// function here and by the HIP kernels:
//   * Each multiply-add that the reference writes as `a*b + c` is evaluated as
//     one fused multiply-add (what nvcc's default --fmad=true contraction
//     does); `__fmaf_rn` calls are std::fma.  The file is built with
//     -ffp-contract=off so that no other contraction happens.
//   * `rsqrtf(x)` is evaluated as 1/sqrt(x) with IEEE sqrt and divide.
//   * Texture fetches (tex2D/tex3D/tex2DLayered with linear filtering,
//     border addressing, unnormalised coordinates) are emulated as: the
//     coordinate c handed to the sampler is formed exactly as the reference
//     forms it (p + 0.5f), then b = c - 0.5f, i = floor(b), a = b - i, and
//     the neighbours outside the array read 0.  Weights are exact unless a
//     caller asks for them quantised to 2^-qbits (the texture unit's 8-bit
//     fractional weights), in which case an error bound is also produced.
//   * Every ray march carries an iteration cap, ceil(1.25*len/step) + 16 and
//     at most 2^25, that the reference lacks.  It binds when the step is below
//     the float resolution of alpha, where the reference would never
//     terminate, and for rays of more than 2^25 samples (on the largest
//     volumes the kernels accept, a step below about 0.43 voxel).
//
// Templates are instantiated for float (the reference precision) and double
// (the FP64 twin used to state errors).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> inline T fma_(T a, T b, T c) { return std::fma(a, b, c); }

// helper_grid.h:23-26 index_to_physical (i*s + o, contracted), :28-31 physical_to_index.
template <class T> inline T i2p(T i, T o, T s) { return fma_(i, s, o); }
template <class T> inline T p2i(T p, T o, T s) { return (p - o) / s; }

// Linear blend used by every interpolator:
// cone_projector_3D_CudaKernel.cu.cc:74-82 writes __fmaf_rn(b, t, __fmaf_rn(a, -t, a)).
template <class T> inline T lerp_(T a, T b, T t) { return fma_(b, t, fma_(a, -t, a)); }

template <class T> inline T quantise(T t, int qbits) {
  if (qbits <= 0) return t;
  const T scale = T(1 << qbits);
  return std::nearbyint(t * scale) / scale;
}

// ---------------------------------------------------------------------------
// Trilinear sampler, zero outside [0,N-1] per axis.
// cone_projector_3D_CudaKernel.cu.cc:42-85 (interp3D): floor, 8 neighbours,
// lerp in z, then y, then x.
// ---------------------------------------------------------------------------
template <class T> struct Vol3 {
  const T *p;
  int W, H, D;
  inline T at(int x, int y, int z) const {
    if (x < 0 || x >= W || y < 0 || y >= H || z < 0 || z >= D) return T(0);
    return p[((size_t)z * H + y) * W + x];
  }
};

template <class T>
inline T interp3(const Vol3<T> &v, T px, T py, T pz, int qbits, T *range) {
  // Points reaching here lie inside the clipped box; anything else (NaN,
  // far outside) has all eight neighbours outside the volume.
  if (!(px > T(-2) && px < T(v.W + 1) && py > T(-2) && py < T(v.H + 1) && pz > T(-2) &&
        pz < T(v.D + 1))) {
    if (range) *range = 0;
    return T(0);
  }
  const T fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
  const int xf = (int)fx, yf = (int)fy, zf = (int)fz;
  const int xc = xf + 1, yc = yf + 1, zc = zf + 1;
  const T p000 = v.at(xf, yf, zf), p001 = v.at(xc, yf, zf);
  const T p010 = v.at(xf, yc, zf), p011 = v.at(xc, yc, zf);
  const T p100 = v.at(xf, yf, zc), p101 = v.at(xc, yf, zc);
  const T p110 = v.at(xf, yc, zc), p111 = v.at(xc, yc, zc);
  const T xd = quantise(px - fx, qbits);
  const T yd = quantise(py - fy, qbits);
  const T zd = quantise(pz - fz, qbits);
  if (range) {
    T lo = std::min({p000, p001, p010, p011, p100, p101, p110, p111});
    T hi = std::max({p000, p001, p010, p011, p100, p101, p110, p111});
    *range = hi - lo;
  }
  const T p00 = lerp_(p000, p100, zd);
  const T p01 = lerp_(p001, p101, zd);
  const T p10 = lerp_(p010, p110, zd);
  const T p11 = lerp_(p011, p111, zd);
  const T p0 = lerp_(p00, p10, yd);
  const T p1 = lerp_(p01, p11, yd);
  return lerp_(p0, p1, xd);
}

// ---------------------------------------------------------------------------
// Bilinear sampler on one detector image / 2D volume [rows][cols], zero
// outside.  Lerp in x (columns), then y: cone_backprojector_3D_CudaKernel.cu.cc:55-82.
// `quirk` reproduces that kernel's bounds test (x > W, y > H, Q11) with the
// linear index running on into the next row / view; reads past the end of
// the whole buffer return 0.
// ---------------------------------------------------------------------------
template <class T> struct Img2 {
  const T *p;    // start of this image
  int W, H;      // cols, rows
  const T *base; // start of the whole buffer (quirk mode)
  size_t total;  // elements in the whole buffer (quirk mode)
  inline T at(int x, int y) const {
    if (x < 0 || x >= W || y < 0 || y >= H) return T(0);
    return p[(size_t)y * W + x];
  }
  inline T at_quirk(int x, int y) const {
    if (x < 0 || x > W || y < 0 || y > H) return T(0);
    const size_t off = (size_t)(p - base) + (size_t)y * W + x;
    return off < total ? base[off] : T(0);
  }
};

template <class T>
inline T interp2(const Img2<T> &im, T px, T py, int qbits, bool quirk, T *range) {
  if (!(px > T(-2) && px < T(im.W + 1) && py > T(-2) && py < T(im.H + 1))) {
    if (range) *range = 0;
    return T(0);
  }
  const T fx = std::floor(px), fy = std::floor(py);
  const int xf = (int)fx, yf = (int)fy, xc = xf + 1, yc = yf + 1;
  T p00, p01, p10, p11;
  if (quirk) {
    p00 = im.at_quirk(xf, yf); p01 = im.at_quirk(xc, yf);
    p10 = im.at_quirk(xf, yc); p11 = im.at_quirk(xc, yc);
  } else {
    p00 = im.at(xf, yf); p01 = im.at(xc, yf);
    p10 = im.at(xf, yc); p11 = im.at(xc, yc);
  }
  const T xd = quantise(px - fx, qbits);
  const T yd = quantise(py - fy, qbits);
  if (range) *range = std::max({p00, p01, p10, p11}) - std::min({p00, p01, p10, p11});
  const T p0 = lerp_(p00, p01, xd);
  const T p1 = lerp_(p10, p11, xd);
  return lerp_(p0, p1, yd);
}

// Linear sampler along one sinogram row (2D backprojection: the row coordinate
// n + 0.5 lands exactly on row n, so only the detector axis interpolates).
// par_backprojector_2D_CudaKernel.cu.cc:54, fan_backprojector_2D_CudaKernel.cu.cc:60.
template <class T> inline T interp1(const T *row, int W, T px, int qbits, T *range) {
  if (!(px > T(-2) && px < T(W + 1))) {
    if (range) *range = 0;
    return T(0);
  }
  const T fx = std::floor(px);
  const int xf = (int)fx, xc = xf + 1;
  const T p0 = (xf >= 0 && xf < W) ? row[xf] : T(0);
  const T p1 = (xc >= 0 && xc < W) ? row[xc] : T(0);
  if (range) *range = std::fabs(p1 - p0);
  return lerp_(p0, p1, quantise(px - fx, qbits));
}

// Iteration cap shared with the HIP kernels.
// Evaluated in float, exactly as the kernels do; at most 2^25 iterations.
template <class T> inline long long march_cap(T amin, T amax, T step) {
  if (!(amax > amin)) return 0;
  const float n = std::ceil(1.25f * ((float)amax - (float)amin) / (float)step) + 16.0f;
  return n < 33554432.0f ? (long long)n : 33554432LL;
}

// ---------------------------------------------------------------------------
// Cone-beam ray march.
// SW: cone_projector_3D_CudaKernel.cu.cc:87-181 (box [0,N], interp at p).
// HW: cone_projector_3D_CudaKernel_hardware_interp.cu.cc:43-142 (box
//     [-0.5,N-0.5], sampler coordinate p + 0.5).
// ---------------------------------------------------------------------------
template <class T> struct Sampler3 {
  const Vol3<T> *v;
  bool hw;
  int qbits;
  inline T operator()(T px, T py, T pz, T *range) const {
    if (hw) {
      const T h = T(0.5);
      return interp3(*v, (px + h) - h, (py + h) - h, (pz + h) - h, qbits, range);
    }
    return interp3(*v, px, py, pz, qbits, range);
  }
};

template <class T>
T cone_ray(const Sampler3<T> &f, const T s[3], const T r[3], T step, T *bound, long long *nsamp) {
  const int N[3] = {f.v->W, f.v->H, f.v->D};
  if (bound) *bound = 0;
  if (nsamp) *nsamp = 0;
  if (!(step > T(0))) return std::numeric_limits<T>::quiet_NaN();  // reference would not terminate
  T amin = T(0), amax = std::numeric_limits<T>::infinity();
  for (int a = 0; a < 3; ++a) {
    if (T(0) != r[a]) {
      const T e0 = f.hw ? T(0) - T(0.5) : T(0);
      const T e1 = f.hw ? T(N[a]) - T(0.5) : T(N[a]);
      const T reci = T(1) / r[a];
      const T a0 = (e0 - s[a]) * reci;
      const T a1 = (e1 - s[a]) * reci;
      if (a == 0) {  // the x slab overwrites the initial [0, inf) (Q18)
        amin = std::fmin(a0, a1);
        amax = std::fmax(a0, a1);
      } else {
        amin = std::fmax(amin, std::fmin(a0, a1));
        amax = std::fmin(amax, std::fmax(a0, a1));
      }
    }
  }
  amin = fma_(step, T(0.5), amin);  // start half a step inside (Q4)
  T pix = T(0);
  T px = 0, py = 0, pz = 0;
  T bnd = T(0), rg = T(0);
  T *rp = bound ? &rg : nullptr;
  long long ns = 0;
  const long long cap = march_cap(amin, amax, step);
  if (amin < amax) {
    px = fma_(amin, r[0], s[0]);
    py = fma_(amin, r[1], s[1]);
    pz = fma_(amin, r[2], s[2]);
    pix = fma_(T(0.5), f(px, py, pz, rp), pix);
    bnd += T(0.5) * rg;
    amin += step;
    ++ns;
  }
  long long it = 0;
  while (amin < amax && it < cap) {
    px = fma_(amin, r[0], s[0]);
    py = fma_(amin, r[1], s[1]);
    pz = fma_(amin, r[2], s[2]);
    pix += f(px, py, pz, rp);
    bnd += rg;
    amin += step;
    ++it;
  }
  ns += it;
  if (ns > 0) ++ns;  // exit sample (counted whether or not the tail runs)
  pix *= step;
  bnd *= step;
  if (pix > T(0)) {  // tail only for a positive sum (Q6)
    pix = fma_(-(T(0.5) * step), f(px, py, pz, nullptr), pix);
    amin -= step;
    const T last = amax - amin;
    T r1 = 0;
    pix = fma_(T(0.5) * last, f(px, py, pz, rp ? &r1 : nullptr), pix);
    px = fma_(amax, r[0], s[0]);
    py = fma_(amax, r[1], s[1]);
    pz = fma_(amax, r[2], s[2]);
    T r2 = 0;
    pix = fma_(T(0.5) * last, f(px, py, pz, rp ? &r2 : nullptr), pix);
    bnd += T(0.5) * std::fabs(last) * (r1 + r2) + T(0.5) * step * r1;
  }
  if (bound) *bound = bnd;
  if (nsamp) *nsamp = ns;
  return pix;
}

// Per-pixel ray direction and output scale:
// cone_projector_3D_CudaKernel.cu.cc:197-221.
template <class T>
void cone_fwd(const T *vol, int D, int H, int W, const T *inv_ar, const T *src, int n, int v0, int v1,
              int Hd, int Wd, const T *spacing_zyx, T step, int hw, int qbits, T *out, T *bound,
              long long *nsamp_total) {
  Vol3<T> v{vol, W, H, D};
  Sampler3<T> smp{&v, hw != 0, qbits};
  const T sp[3] = {spacing_zyx[2], spacing_zyx[1], spacing_zyx[0]};
  long long total = 0;
  (void)n;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
  for (long long vr = (long long)v0 * Hd; vr < (long long)v1 * Hd; ++vr) {
    const int view = (int)(vr / Hd), row = (int)(vr % Hd);
    const T *m = inv_ar + (size_t)view * 9;
    const T s[3] = {src[view * 3 + 0], src[view * 3 + 1], src[view * 3 + 2]};
    for (int col = 0; col < Wd; ++col) {
      const T u = T(col), w = T(row);
      T r[3];
      r[0] = fma_(u, m[0], fma_(w, m[1], m[2]));
      r[1] = fma_(u, m[3], fma_(w, m[4], m[5]));
      r[2] = fma_(u, m[6], fma_(w, m[7], m[8]));
      const T nrm = T(1) / std::sqrt(fma_(r[2], r[2], fma_(r[1], r[1], r[0] * r[0])));
      r[0] *= nrm; r[1] *= nrm; r[2] *= nrm;
      T b = 0;
      long long ns = 0;
      T pix = cone_ray(smp, s, r, step, bound ? &b : nullptr, &ns);
      const T a = r[0] * sp[0], bb = r[1] * sp[1], c = r[2] * sp[2];
      const T scale = std::sqrt(fma_(c, c, fma_(bb, bb, a * a)));
      pix *= scale;
      const size_t o = ((size_t)(view - v0) * Hd + row) * Wd + col;
      out[o] = pix;
      if (bound) bound[o] = b * scale;
      total += ns;
    }
  }
  if (nsamp_total) *nsamp_total = total;
}

// ---------------------------------------------------------------------------
// Cone-beam voxel-driven backprojection.
// SW: cone_backprojector_3D_CudaKernel.cu.cc:44-53 (map), :84-118.
// HW: cone_backprojector_3D_CudaKernel_hardware_interp.cu.cc:55-87 (sampler
//     coordinate u + 0.5, v + 0.5 on layer n).
// ---------------------------------------------------------------------------
template <class T>
void cone_bp(const T *sino, int n, int Hd, int Wd, const T *P, int D, int H, int W,
             const T *spacing_zyx, const T *origin_zyx, T mult, int hw, int qbits, int quirk,
             int z0, int z1, T *out, T *bound) {
  const T sx = spacing_zyx[2], sy = spacing_zyx[1], sz = spacing_zyx[0];
  const T ox = origin_zyx[2], oy = origin_zyx[1], oz = origin_zyx[0];
  const size_t total = (size_t)n * Hd * Wd;
  (void)D;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int k = z0; k < z1; ++k) {
    for (int j = 0; j < H; ++j) {
      for (int i = 0; i < W; ++i) {
        const T x = i2p(T(i), ox, sx), y = i2p(T(j), oy, sy), z = i2p(T(k), oz, sz);
        T val = T(0), bnd = T(0);
        for (int v = 0; v < n; ++v) {
          const T *m = P + (size_t)v * 12;
          T u = m[0] * x; u = fma_(m[1], y, u); u = fma_(m[2], z, u); u = u + m[3];
          T w = m[4] * x; w = fma_(m[5], y, w); w = fma_(m[6], z, w); w = w + m[7];
          T q = m[8] * x; q = fma_(m[9], y, q); q = fma_(m[10], z, q); q = q + m[11];
          q = T(1) / q;
          u *= q;
          w *= q;
          Img2<T> im{sino + (size_t)v * Hd * Wd, Wd, Hd, sino, total};
          T rg = 0;
          T f;
          if (hw) {
            const T h = T(0.5);
            f = interp2(im, (u + h) - h, (w + h) - h, qbits, false, bound ? &rg : nullptr);
          } else {
            f = interp2(im, u, w, qbits, quirk != 0, bound ? &rg : nullptr);
          }
          val = fma_(f * q, q, val);
          bnd += rg * q * q;
        }
        const size_t o = ((size_t)(k - z0) * H + j) * W + i;
        out[o] = val * mult;
        if (bound) bound[o] = T(2) * bnd * std::fabs(mult);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 2D ray march in physical units (texture sampler on the volume).
// par_projector_2D_CudaKernel.cu.cc:28-102 (identical copy in
// fan_projector_2D_CudaKernel.cu.cc:27-102).  Box edges carry the -0.5 mm
// offset (Q8); the march starts exactly at the entry point (Q4).
// ---------------------------------------------------------------------------
template <class T> struct Sampler2 {
  Img2<T> im;
  T ox, oy, sx, sy;
  int qbits;
  inline T operator()(T px, T py, T *range) const {
    const T h = T(0.5);
    const T cx = p2i(px, ox, sx) + h, cy = p2i(py, oy, sy) + h;
    return interp2(im, cx - h, cy - h, qbits, false, range);
  }
};

template <class T>
T march2d(const Sampler2<T> &f, const T s[2], const T r[2], T step, int W, int H, T *bound,
          long long *nsamp) {
  if (bound) *bound = 0;
  if (nsamp) *nsamp = 0;
  if (!(step > T(0))) return std::numeric_limits<T>::quiet_NaN();
  T amin = T(0), amax = std::numeric_limits<T>::infinity();
  if (T(0) != r[0]) {
    const T e0 = i2p(T(0), f.ox, f.sx) - T(0.5);
    const T e1 = i2p(T(W), f.ox, f.sx) - T(0.5);
    const T reci = T(1) / r[0];
    const T a0 = (e0 - s[0]) * reci, a1 = (e1 - s[0]) * reci;
    amin = std::fmin(a0, a1);
    amax = std::fmax(a0, a1);
  }
  if (T(0) != r[1]) {
    const T e0 = i2p(T(0), f.oy, f.sy) - T(0.5);
    const T e1 = i2p(T(H), f.oy, f.sy) - T(0.5);
    const T reci = T(1) / r[1];
    const T a0 = (e0 - s[1]) * reci, a1 = (e1 - s[1]) * reci;
    amin = std::fmax(amin, std::fmin(a0, a1));
    amax = std::fmin(amax, std::fmax(a0, a1));
  }
  T pix = 0, px = 0, py = 0, bnd = 0, rg = 0;
  T *rp = bound ? &rg : nullptr;
  long long ns = 0;
  const long long cap = march_cap(amin, amax, step);
  if (amin < amax) {
    px = fma_(amin, r[0], s[0]);
    py = fma_(amin, r[1], s[1]);
    pix = fma_(T(0.5), f(px, py, rp), pix);
    bnd += T(0.5) * rg;
    amin += step;
    ++ns;
  }
  long long it = 0;
  while (amin < amax && it < cap) {
    px = fma_(amin, r[0], s[0]);
    py = fma_(amin, r[1], s[1]);
    pix += f(px, py, rp);
    bnd += rg;
    amin += step;
    ++it;
  }
  ns += it;
  if (ns > 0) ++ns;
  pix *= step;
  bnd *= step;
  if (pix > T(0)) {
    pix = fma_(-(T(0.5) * step), f(px, py, nullptr), pix);
    amin -= step;
    const T last = amax - amin;
    T r1 = 0, r2 = 0;
    pix = fma_(T(0.5) * last, f(px, py, rp ? &r1 : nullptr), pix);
    px = fma_(amax, r[0], s[0]);
    py = fma_(amax, r[1], s[1]);
    pix = fma_(T(0.5) * last, f(px, py, rp ? &r2 : nullptr), pix);
    bnd += T(0.5) * std::fabs(last) * (r1 + r2) + T(0.5) * step * r1;
  }
  if (bound) *bound = bnd;
  if (nsamp) *nsamp = ns;
  return pix;
}

// Parallel: par_projector_2D_CudaKernel.cu.cc:104-147; fan: fan_projector_2D_CudaKernel.cu.cc:104-149.
template <class T>
void proj2d(bool fan, const T *vol, int H, int W, const T *rays, int n, int det,
            const T *vol_origin_yx, const T *vol_spacing_yx, T det_origin, T det_spacing, T sid,
            T sdd, T step_factor, int qbits, T *out, T *bound, long long *nsamp_total) {
  const T sx = vol_spacing_yx[1], sy = vol_spacing_yx[0];
  const T ox = vol_origin_yx[1], oy = vol_origin_yx[0];
  Sampler2<T> smp{Img2<T>{vol, W, H, vol, (size_t)W * H}, ox, oy, sx, sy, qbits};
  const T step = step_factor * std::fmin(sx, sy);
  T psid = sid;
  if (!fan) {
    // Virtual source distance, par_projector_2D_CudaKernel.cu.cc:120.
    const T a = ((T(W) * sx) * T(W)) * sx;
    psid = std::sqrt(fma_((T(H) * sy) * T(H), sy, a)) * T(1.2);
  }
  long long total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
  for (long long idx = 0; idx < (long long)n * det; ++idx) {
    const int view = (int)(idx / det), d = (int)(idx % det);
    const T cr[2] = {rays[view * 2 + 0], rays[view * 2 + 1]};
    const T uv[2] = {-cr[1], cr[0]};
    const T u = i2p(T(d), det_origin, det_spacing);
    T s[2], r[2];
    if (!fan) {
      s[0] = fma_(uv[0], u, cr[0] * (-psid));
      s[1] = fma_(uv[1], u, cr[1] * (-psid));
      r[0] = cr[0];
      r[1] = cr[1];
    } else {
      s[0] = cr[0] * (-sid);
      s[1] = cr[1] * (-sid);
      const T dx = fma_(uv[0], u, fma_(cr[0], sdd, s[0]));
      const T dy = fma_(uv[1], u, fma_(cr[1], sdd, s[1]));
      const T ex = dx - s[0], ey = dy - s[1];
      const T nrm = T(1) / std::sqrt(fma_(ey, ey, ex * ex));
      r[0] = ex * nrm;
      r[1] = ey * nrm;
    }
    T b = 0;
    long long ns = 0;
    T pix = march2d(smp, s, r, step, W, H, bound ? &b : nullptr, &ns);
    const T a = r[0] * sx, c = r[1] * sy;
    const T scale = std::sqrt(fma_(c, c, a * a));
    out[idx] = pix * scale;
    if (bound) bound[idx] = b * scale;
    total += ns;
  }
  if (nsamp_total) *nsamp_total = total;
}

// 2D line intersection, helper_geometry_gpu.h:28-44 (NaN when |den| < 1e-6, Q15).
template <class T>
inline void isect2(const T p1[2], const T p2[2], const T p3[2], const T p4[2], T out[2]) {
  const T a = p1[0] - p2[0], b = p3[1] - p4[1], c = p1[1] - p2[1], d = p3[0] - p4[0];
  const T den = fma_(-c, d, a * b);
  if (den < T(0.000001) && den > T(-0.000001)) {
    out[0] = out[1] = std::numeric_limits<T>::quiet_NaN();
    return;
  }
  const T e = fma_(-p1[1], p2[0], p1[0] * p2[1]);
  const T g = fma_(-p3[1], p4[0], p3[0] * p4[1]);
  T x = fma_(-a, g, e * d);
  T y = fma_(-c, g, e * b);
  out[0] = x / den;
  out[1] = y / den;
}

// Parallel: par_backprojector_2D_CudaKernel.cu.cc:28-61; fan: fan_backprojector_2D_CudaKernel.cu.cc:29-67.
template <class T>
void back2d(bool fan, const T *sino, int n, int det, const T *rays, int H, int W,
            const T *vol_origin_yx, const T *vol_spacing_yx, T det_origin, T det_spacing, T sid,
            T sdd, int qbits, T *out, T *bound) {
  const T sx = vol_spacing_yx[1], sy = vol_spacing_yx[0];
  const T ox = vol_origin_yx[1], oy = vol_origin_yx[0];
  const T pi = T(3.14159265359);
  const T h = T(0.5);
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      const T pc[2] = {i2p(T(x), ox, sx), i2p(T(y), oy, sy)};
      T val = 0, bnd = 0;
      for (int v = 0; v < n; ++v) {
        const T cr[2] = {rays[v * 2 + 0], rays[v * 2 + 1]};
        const T dv[2] = {-cr[1], cr[0]};
        const T *row = sino + (size_t)v * det;
        T rg = 0;
        if (!fan) {
          const T s = fma_(pc[1], dv[1], pc[0] * dv[0]);
          const T si = p2i(s, det_origin, det_spacing);
          val += interp1(row, det, (si + h) - h, qbits, bound ? &rg : nullptr);
          bnd += rg;
        } else {
          const T src[2] = {cr[0] * (-sid), cr[1] * (-sid)};
          const T cp[2] = {fma_(cr[0], sdd, src[0]), fma_(cr[1], sdd, src[1])};
          const T cp2[2] = {cp[0] + dv[0], cp[1] + dv[1]};
          T is[2];
          isect2(pc, src, cp, cp2, is);
          const T ex = pc[0] - src[0], ey = pc[1] - src[1];
          const T dw = T(1) / std::sqrt(fma_(ey, ey, ex * ex));
          const T s = fma_(is[1], dv[1], is[0] * dv[0]);
          const T si = p2i(s, det_origin, det_spacing);
          const T f = interp1(row, det, (si + h) - h, qbits, bound ? &rg : nullptr);
          val = fma_(f * dw, dw, val);
          bnd += rg * dw * dw;
        }
      }
      const size_t o = (size_t)y * W + x;
      if (!fan) {
        out[o] = (pi * val) / T(n);
        if (bound) bound[o] = (pi * bnd) / T(n);
      } else {
        out[o] = (((sid * sdd) * pi) * val) / T(n);
        if (bound) bound[o] = std::fabs(((sid * sdd) * pi) * bnd) / T(n);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Per-view cone geometry in FP64, closed form.
// Replaces cone_projector_3D_OPKernel.cc:96-182 and helper_geometry_cpu.h:30-43
// (Eigen JacobiSVD null space, normalised so C(3) = 1): for P = [M | p4] the
// right null vector with C(3) = 1 is C = -M^{-1} p4.  M^{-1} = adj(M) / det(M).
//   src_idx = (C - origin) / spacing           (cone_projector_3D_OPKernel.cc:167-169)
//   inv_AR  = diag(1/spacing) * M^{-1}         (cone_projector_3D_OPKernel.cc:171)
// The same operation order is used by the HIP setup kernel.
// Returns 0, or 1 when |det M| <= 1e-300 or non-finite results appear.
// ---------------------------------------------------------------------------
int cone_setup_one(const float *Pf, const float *spacing_zyx, const float *origin_zyx,
                   double inv[9], double src[3]) {
  double m[12];
  for (int i = 0; i < 12; ++i) m[i] = (double)Pf[i];
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9],
               k = m[10];
  const double A = e * k - f * h, B = f * g - d * k, C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  if (!(std::fabs(det) > 1e-300) || !std::isfinite(det)) return 1;
  double adj[9] = {A, c * h - b * k, b * f - c * e, B, a * k - c * g, c * d - a * f,
                   C, b * g - a * h, a * e - b * d};
  double mi[9];
  for (int i = 0; i < 9; ++i) mi[i] = adj[i] / det;
  const double p4[3] = {m[3], m[7], m[11]};
  const double sp[3] = {(double)spacing_zyx[2], (double)spacing_zyx[1], (double)spacing_zyx[0]};
  const double og[3] = {(double)origin_zyx[2], (double)origin_zyx[1], (double)origin_zyx[0]};
  for (int r = 0; r < 3; ++r) {
    const double cr = -(mi[r * 3 + 0] * p4[0] + mi[r * 3 + 1] * p4[1] + mi[r * 3 + 2] * p4[2]);
    src[r] = (cr - og[r]) / sp[r];
    for (int cc = 0; cc < 3; ++cc) inv[r * 3 + cc] = mi[r * 3 + cc] / sp[r];
  }
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(inv[i])) return 1;
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(src[i])) return 1;
  return 0;
}

}  // namespace

extern "C" {

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_cone_setup(const float *P, int n, const float *spacing_zyx, const float *origin_zyx,
                   float *inv_ar, float *src, double *inv_ar64, double *src64) {
  int bad = 0;
  for (int v = 0; v < n; ++v) {
    double inv[9], s[3];
    if (cone_setup_one(P + (size_t)v * 12, spacing_zyx, origin_zyx, inv, s)) {
      bad = 1;
      for (int i = 0; i < 9; ++i) inv[i] = std::numeric_limits<double>::quiet_NaN();
      for (int i = 0; i < 3; ++i) s[i] = std::numeric_limits<double>::quiet_NaN();
    }
    for (int i = 0; i < 9; ++i) {
      if (inv_ar) inv_ar[v * 9 + i] = (float)inv[i];
      if (inv_ar64) inv_ar64[v * 9 + i] = inv[i];
    }
    for (int i = 0; i < 3; ++i) {
      if (src) src[v * 3 + i] = (float)s[i];
      if (src64) src64[v * 3 + i] = s[i];
    }
  }
  return bad;
}

void orc_cone_fwd_f32(const float *vol, int D, int H, int W, const float *inv_ar, const float *src,
                      int n, int v0, int v1, int Hd, int Wd, const float *spacing_zyx, float step,
                      int hw, int qbits, float *out, float *bound, long long *nsamp) {
  cone_fwd<float>(vol, D, H, W, inv_ar, src, n, v0, v1, Hd, Wd, spacing_zyx, step, hw, qbits, out,
                  bound, nsamp);
}

void orc_cone_fwd_f64(const double *vol, int D, int H, int W, const double *inv_ar,
                      const double *src, int n, int v0, int v1, int Hd, int Wd,
                      const double *spacing_zyx, double step, int hw, double *out,
                      long long *nsamp) {
  cone_fwd<double>(vol, D, H, W, inv_ar, src, n, v0, v1, Hd, Wd, spacing_zyx, step, hw, 0, out,
                   nullptr, nsamp);
}

void orc_cone_bp_f32(const float *sino, int n, int Hd, int Wd, const float *P, int D, int H, int W,
                     const float *spacing_zyx, const float *origin_zyx, float mult, int hw,
                     int qbits, int quirk, int z0, int z1, float *out, float *bound) {
  cone_bp<float>(sino, n, Hd, Wd, P, D, H, W, spacing_zyx, origin_zyx, mult, hw, qbits, quirk, z0,
                 z1, out, bound);
}

void orc_cone_bp_f64(const double *sino, int n, int Hd, int Wd, const double *P, int D, int H,
                     int W, const double *spacing_zyx, const double *origin_zyx, double mult,
                     int hw, int z0, int z1, double *out) {
  cone_bp<double>(sino, n, Hd, Wd, P, D, H, W, spacing_zyx, origin_zyx, mult, hw, 0, 0, z0, z1,
                  out, nullptr);
}

void orc_proj2d_f32(int fan, const float *vol, int H, int W, const float *rays, int n, int det,
                    const float *vol_origin_yx, const float *vol_spacing_yx, float det_origin,
                    float det_spacing, float sid, float sdd, float step_factor, int qbits,
                    float *out, float *bound, long long *nsamp) {
  proj2d<float>(fan != 0, vol, H, W, rays, n, det, vol_origin_yx, vol_spacing_yx, det_origin,
                det_spacing, sid, sdd, step_factor, qbits, out, bound, nsamp);
}

void orc_proj2d_f64(int fan, const double *vol, int H, int W, const double *rays, int n, int det,
                    const double *vol_origin_yx, const double *vol_spacing_yx, double det_origin,
                    double det_spacing, double sid, double sdd, double step_factor, double *out,
                    long long *nsamp) {
  proj2d<double>(fan != 0, vol, H, W, rays, n, det, vol_origin_yx, vol_spacing_yx, det_origin,
                 det_spacing, sid, sdd, step_factor, 0, out, nullptr, nsamp);
}

void orc_back2d_f32(int fan, const float *sino, int n, int det, const float *rays, int H, int W,
                    const float *vol_origin_yx, const float *vol_spacing_yx, float det_origin,
                    float det_spacing, float sid, float sdd, int qbits, float *out, float *bound) {
  back2d<float>(fan != 0, sino, n, det, rays, H, W, vol_origin_yx, vol_spacing_yx, det_origin,
                det_spacing, sid, sdd, qbits, out, bound);
}

void orc_back2d_f64(int fan, const double *sino, int n, int det, const double *rays, int H, int W,
                    const double *vol_origin_yx, const double *vol_spacing_yx, double det_origin,
                    double det_spacing, double sid, double sdd, double *out) {
  back2d<double>(fan != 0, sino, n, det, rays, H, W, vol_origin_yx, vol_spacing_yx, det_origin,
                 det_spacing, sid, sdd, 0, out, nullptr);
}

}  // extern "C"
