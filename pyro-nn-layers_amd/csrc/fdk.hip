// FDK / FBP pre-processing: cosine weighting and Ram-Lak ramp filtering of
// detector rows.  Not part of the reference (its FDK weights and filters live in
// the companion pyronn package); the formulas follow Kak & Slaney, "Principles
// of Computerized Tomographic Imaging", ch. 3:
//   cosine weight  c(u, v) = SDD / sqrt(SDD^2 + a^2 + b^2), (a, b) the detector
//                  coordinates in mm from the principal point (b = 0 for fan);
//   ramp kernel    h(0) = 1 / (4 tau^2), h(k odd) = -1 / (k^2 pi^2 tau^2),
//                  h(k even, k != 0) = 0, tau the sample spacing;
//   filtering      Q_c = tau * sum_j R'_j h(c - j), a linear (zero-padded)
//                  convolution, i.e. the product of each row with the symmetric
//                  Toeplitz matrix T[j][c] = tau * h(c - j).
// The row filtering is a plain GEMM [rows x W_d] x [W_d x W_d] and runs in
// rocBLAS; the weighting, the matrix build and the storage conversions are
// HIP kernels here.
#include <rocblas/rocblas.h>

#include <mutex>
#include <unordered_map>

#include "common.h"

namespace pyronn {

__global__ void ramp_matrix_kernel(float *__restrict__ T, int W, double tau,
                                   const float *__restrict__ colw) {
  const long long total = (long long)W * W;
  const double pi = 3.14159265358979323846;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(t / W), c = (int)(t - (long long)j * W);
    const int k = c - j;
    double h;
    if (k == 0) h = 1.0 / (4.0 * tau * tau);
    else if (k & 1) h = -1.0 / ((double)k * (double)k * pi * pi * tau * tau);
    else h = 0.0;
    T[t] = (float)(tau * h * (colw ? (double)colw[c] : 1.0));
  }
}

__global__ void cosine_weights_kernel(float *__restrict__ w, int Hd, int Wd, double du, double dv,
                                      double u0, double v0, double sdd) {
  const long long total = (long long)Hd * Wd;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(t / Wd), c = (int)(t - (long long)r * Wd);
    const double a = ((double)c - u0) * du, b = ((double)r - v0) * dv;
    w[t] = (float)(sdd / sqrt(sdd * sdd + a * a + b * b));
  }
}

// out[img][r][c] = in[img][r][c] * w[r][c]  (w == nullptr: plain conversion to f32)
template <typename TS>
__global__ void weight_kernel(const TS *__restrict__ in, const float *__restrict__ w,
                              float *__restrict__ out, long long per_image, long long images) {
  // x: pixel within an image, y: image (strided), so the weight index needs no division
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_image) return;
  const float wi = w ? w[i] : 1.0f;
  for (long long img = blockIdx.y; img < images; img += gridDim.y) {
    const long long t = img * per_image + i;
    const float x = to_f32(in[t]);
    out[t] = w ? x * wi : x;
  }
}

__device__ __forceinline__ void store_as(float x, float *p) { *p = x; }
__device__ __forceinline__ void store_as(float x, bf16s *p) {
  // round to nearest even (the bf16 rounding torch uses); NaN stays NaN
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    p->bits = (uint16_t)((u >> 16) | 0x40u);
    return;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  p->bits = (uint16_t)(u >> 16);
}
__device__ __forceinline__ void store_as(float x, _Float16 *p) { *p = (_Float16)x; }

template <typename TO>
__global__ void convert_kernel(const float *__restrict__ in, TO *__restrict__ out,
                               long long total) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x)
    store_as(in[t], out + t);
}

static unsigned grid_for(long long total) {
  long long b = (total + 255) / 256;
  return (unsigned)(b > 65536 * 4 ? 65536 * 4 : (b < 1 ? 1 : b));
}

// One rocBLAS handle per (thread, device); rocBLAS handles are not thread safe.
static rocblas_handle blas_handle(int &err) {
  static thread_local std::unordered_map<int, rocblas_handle> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    err = 1;
    return nullptr;
  }
  auto it = handles.find(dev);
  if (it != handles.end()) return it->second;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) {
    err = 1;
    return nullptr;
  }
  handles[dev] = h;
  return h;
}


// ----------------------------------------------------------------- FFT ramp filter
// The same filter as the Toeplitz GEMM, evaluated as a circular convolution of length
// L = 2^m >= 2 W - 1 (zero padding makes it the linear one, Kak & Slaney ch. 3):
// Q = IFFT(H . FFT(x)), H[k] = h(0) + 2 sum_{n=1}^{W-1} h(n) cos(2 pi k n / L) (real:
// the kernel is even).  One workgroup filters two rows of one image at once as the
// real and imaginary parts of one complex sequence (H is real, so the two stay
// separate), entirely in LDS: one read of the input and one write of the output.
// Rows are paired within an image (rows 2p and 2p + 1; an odd last row is paired with
// zeros), so a row's result depends on its image alone, not on how many images one
// call filters.
//
// FFT: Stockham autosort, radix 8 passes after one radix-2 or radix-4 pass, L / 8
// lanes each holding 8 values of every pass (one radix-8, two radix-4 or four radix-2
// butterflies).  The inverse transform is the forward one on conj(H . X) (H folds in
// 1 / L and tau), conjugated at the end.  The forward transform's last pass leaves the
// spectrum in natural order in exactly the registers that the inverse transform's first
// pass (same radix, no twiddles) reads, so the spectral product needs no LDS round trip.
// The kernel is instantiated per transform size, so every radix, stride and LDS offset is
// a compile-time constant (one address computation per butterfly instead of one per
// value: measured 0.39 -> 0.27 ms per 90-view 1024 x 768 range).
//
// Memory side: W <= L / 2, so the second half of the padded sequence is zero and only the
// first half of the result is kept.  Each lane moves four consecutive samples of each row
// (one vector load or store per row when W % 4 == 0), through LDS in and out; the first
// pass skips its zero inputs and the last pass computes only the outputs that are kept.
// The table holds the spectrum in the order the lanes read it (two 16-byte loads per
// lane) and, per pass, the twiddle pairs (w, w^2) of each butterfly (one 16-byte load).
constexpr int kFftMinLog2 = 4, kFftMaxLog2 = 13;  // 16 .. 8192 points, 2 .. 1024 lanes

static int fft_log2(int det_w) {
  int m = kFftMinLog2;
  while ((1LL << m) < 2LL * det_w - 1) ++m;
  return m;
}

// Pass plan of a transform of 2^m points (host and device).  Forward pass i has radix
// fft_radix(m, i); inverse pass k has the forward radices in reverse order.  Ns of a pass
// is the product of the radices before it.
constexpr int fft_passes(int m) { return (m + 2) / 3; }
constexpr int fft_radix(int m, int i) { return (m % 3 == 0 || i > 0) ? 8 : (m % 3 == 1 ? 2 : 4); }
constexpr int fft_ns(int m, int i) {
  int n = 1;
  for (int k = 0; k < i; ++k) n *= fft_radix(m, k);
  return n;
}
constexpr int ifft_radix(int m, int k) { return fft_radix(m, fft_passes(m) - 1 - k); }
constexpr int ifft_ns(int m, int k) {
  int n = 1;
  for (int i = 0; i < k; ++i) n *= ifft_radix(m, i);
  return n;
}
// Twiddle-pair table: Ns entries for each pass with Ns > 1, forward passes 1 .. P-1 then
// inverse passes 1 .. P-1; these give the first entry of each pass and the total.
constexpr int fft_tw_fwd(int m, int i) {
  int o = 0;
  for (int k = 1; k < i; ++k) o += fft_ns(m, k);
  return o;
}
constexpr int fft_tw_inv(int m, int k) {
  int o = fft_tw_fwd(m, fft_passes(m));
  for (int i = 1; i < k; ++i) o += ifft_ns(m, i);
  return o;
}
constexpr int fft_tw_entries(int m) { return fft_tw_inv(m, fft_passes(m)); }

struct cpx {
  float x, y;
};
__device__ __forceinline__ cpx cadd(cpx a, cpx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cpx cmul(cpx a, cpx b) {
  return {fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ cpx mul_mi(cpx a) { return {a.y, -a.x}; }  // a * (-i)

__device__ __forceinline__ void dft2(cpx &a, cpx &b) {
  const cpx t = a;
  a = cadd(t, b);
  b = csub(t, b);
}
// in place: v0..v3 -> X0..X3, X_k = sum_n v_n e^{-2 pi i k n / 4}
__device__ __forceinline__ void dft4(cpx &v0, cpx &v1, cpx &v2, cpx &v3) {
  const cpx a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = mul_mi(csub(v1, v3));
  v0 = cadd(a0, a2);
  v2 = csub(a0, a2);
  v1 = cadd(a1, a3);
  v3 = csub(a1, a3);
}
// dft4 of (v0, v1, 0, 0)
__device__ __forceinline__ void dft4_half(cpx v0, cpx v1, cpx &x0, cpx &x1, cpx &x2, cpx &x3) {
  const cpx a3 = mul_mi(v1);
  x0 = cadd(v0, v1);
  x2 = csub(v0, v1);
  x1 = cadd(v0, a3);
  x3 = csub(v0, a3);
}
// the last stage of dft8: the dft4s of the even (e) and odd (o) inputs combined
__device__ __forceinline__ void dft8_combine(cpx e0, cpx e1, cpx e2, cpx e3, cpx o0, cpx o1,
                                             cpx o2, cpx o3, cpx *v) {
  const float s = 0.70710678118654752f;  // e^{-2 pi i k / 8}, k = 1, 2, 3 applied to o1..o3
  o1 = {(o1.x + o1.y) * s, (o1.y - o1.x) * s};
  o2 = mul_mi(o2);
  o3 = {(o3.y - o3.x) * s, -(o3.x + o3.y) * s};
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}
__device__ __forceinline__ void dft8(cpx *v) {
  cpx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cpx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  dft8_combine(e0, e1, e2, e3, o0, o1, o2, o3, v);
}

// The first forward pass (Ns = 1, no twiddles) when inputs r >= R / 2 of every butterfly,
// the zero padding, are zero: v[b R + r] for r < R / 2 hold the data.
template <int R>
__device__ __forceinline__ void fft_first_pass_half(cpx *v) {
  if constexpr (R == 8) {
    cpx e0, e1, e2, e3, o0, o1, o2, o3;
    dft4_half(v[0], v[2], e0, e1, e2, e3);
    dft4_half(v[1], v[3], o0, o1, o2, o3);
    dft8_combine(e0, e1, e2, e3, o0, o1, o2, o3, v);
  } else if constexpr (R == 4) {
    dft4_half(v[0], v[1], v[0], v[1], v[2], v[3]);
    dft4_half(v[4], v[5], v[4], v[5], v[6], v[7]);
  } else {
    v[1] = v[0];
    v[3] = v[2];
    v[5] = v[4];
    v[7] = v[6];
  }
}

template <int M>
struct FftPlan {
  static constexpr int L = 1 << M, T = L >> 3, passes = fft_passes(M);
  static constexpr int radix(int i) { return fft_radix(M, i); }
  static constexpr int ns(int i) { return fft_ns(M, i); }
};

// LDS slot (one complex value, 8 bytes) of sequence position p >= 0: one pad slot per
// kLdsRow = 32, which spreads the strided stores of the early passes over the banks (one
// per 16 measured the same, one per 64 slower).
constexpr int kLdsRow = 32;
__device__ __forceinline__ int pad32(int p) { return p + (int)((unsigned)p / (unsigned)kLdsRow); }

// Butterfly b (< 8 / R) of a lane in a pass of radix R has index j = lane + b L / 8; its
// values v[b R + r] come from x[j + r L / R] and go to x[(j / Ns) Ns R + (j % Ns) + r Ns].
template <int R, int NS, int L>
__device__ __forceinline__ void fft_twiddle_dft(cpx *v, int lane, const float4 *__restrict__ tw) {
  constexpr int T = L / 8;
  if constexpr (NS > 1) {
#pragma unroll
    for (int b = 0; b < 8 / R; ++b) {
      // w^r, w = e^{-2 pi i k / (Ns R)}: w and w^2 from the pass's table (one entry per k),
      // higher powers as products of at most two table values or one earlier product
      const int k = (lane + b * T) & (NS - 1);
      const float4 t = tw[k];
      const cpx w1 = {t.x, t.y};
      v[b * R + 1] = cmul(v[b * R + 1], w1);
      if constexpr (R >= 4) {
        const cpx w2 = {t.z, t.w}, w3 = cmul(w2, w1);
        v[b * R + 2] = cmul(v[b * R + 2], w2);
        v[b * R + 3] = cmul(v[b * R + 3], w3);
        if constexpr (R == 8) {
          const cpx w4 = cmul(w2, w2);
          v[b * R + 4] = cmul(v[b * R + 4], w4);
          v[b * R + 5] = cmul(v[b * R + 5], cmul(w4, w1));
          v[b * R + 6] = cmul(v[b * R + 6], cmul(w4, w2));
          v[b * R + 7] = cmul(v[b * R + 7], cmul(w4, w3));
        }
      }
    }
  }
  if constexpr (R == 8) {
    dft8(v);
  } else if constexpr (R == 4) {
    dft4(v[0], v[1], v[2], v[3]);
    dft4(v[4], v[5], v[6], v[7]);
  } else {
    dft2(v[0], v[1]);
    dft2(v[2], v[3]);
    dft2(v[4], v[5]);
    dft2(v[6], v[7]);
  }
}

// Padded LDS slot of position base + r * NS.  Linear in r (one address per butterfly,
// immediate offsets) when NS is a multiple of 32 or when R NS <= 32 (then base % 32 + r NS
// never crosses a 32-word row).
template <int R, int NS>
__device__ __forceinline__ int store_word(int base, int r) {
  if constexpr (NS % kLdsRow == 0) return pad32(base) + r * (NS + NS / kLdsRow);
  else if constexpr (R * NS <= kLdsRow) return pad32(base) + r * NS;
  else return pad32(base + r * NS);
}

template <int R, int NS, int L>
__device__ __forceinline__ void fft_store(const cpx *v, cpx *buf, int lane) {
  constexpr int T = L / 8;
#pragma unroll
  for (int b = 0; b < 8 / R; ++b) {
    const int j = lane + b * T;
    const int base = (j & ~(NS - 1)) * R + (j & (NS - 1));
#pragma unroll
    for (int r = 0; r < R; ++r) {
      buf[store_word<R, NS>(base, r)] = v[b * R + r];
    }
  }
}

template <int R, int L>
__device__ __forceinline__ void fft_load(cpx *v, const cpx *buf, int lane) {
  constexpr int T = L / 8, S = L / R;
#pragma unroll
  for (int b = 0; b < 8 / R; ++b) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int a = (T % kLdsRow == 0 && S % kLdsRow == 0)
                        ? pad32(lane) + (b * T + r * S) / kLdsRow * (kLdsRow + 1)
                        : pad32(lane + b * T + r * S);
      v[b * R + r] = buf[a];
    }
  }
}

// Natural-order position of value q of a lane before a first pass / after a last pass of
// radix R: j + r L / R with j = lane + b L / 8, q = b R + r.
template <int R, int L>
__device__ __forceinline__ int fft_natural(int q, int lane) {
  return lane + (q / R) * (L / 8) + (q % R) * (L / R);
}

// Forward passes 1 .. P-1 (pass 0 is done by the caller), each preceded by the LDS
// transposition out of the previous pass.
template <int M, int I>
__device__ __forceinline__ void fft_forward_passes(cpx *v, cpx *buf, int lane,
                                                   const float4 *__restrict__ tw) {
  using F = FftPlan<M>;
  if constexpr (I < F::passes) {
    __syncthreads();
    fft_store<F::radix(I - 1), F::ns(I - 1), F::L>(v, buf, lane);
    __syncthreads();
    fft_load<F::radix(I), F::L>(v, buf, lane);
    fft_twiddle_dft<F::radix(I), F::ns(I), F::L>(v, lane, tw + fft_tw_fwd(M, I));
    fft_forward_passes<M, I + 1>(v, buf, lane, tw);
  }
}

// Inverse transform: the forward passes in reverse radix order; pass K of the inverse has
// radix radix(P - 1 - K) and Ns the product of the inverse radices before it.
template <int M>
struct FftInverse {
  static constexpr int radix(int k) { return ifft_radix(M, k); }
  static constexpr int ns(int k) { return ifft_ns(M, k); }
};

template <int M, int K>
__device__ __forceinline__ void fft_inverse_passes(cpx *v, cpx *buf, int lane,
                                                   const float4 *__restrict__ tw) {
  using F = FftPlan<M>;
  using I = FftInverse<M>;
  if constexpr (K < F::passes) {
    __syncthreads();
    fft_store<I::radix(K - 1), I::ns(K - 1), F::L>(v, buf, lane);
    __syncthreads();
    fft_load<I::radix(K), F::L>(v, buf, lane);
    fft_twiddle_dft<I::radix(K), I::ns(K), F::L>(v, lane, tw + fft_tw_inv(M, K));
    fft_inverse_passes<M, K + 1>(v, buf, lane, tw);
  }
}

// Four consecutive samples c0 .. c0 + 3 of a row, times the weights (if any), zero from W
// on.  vec: W % 4 == 0 and the row and weights are aligned for one vector load each.
template <typename T>
struct alignas(4 * sizeof(T)) pack4 {
  T e[4];
};

template <typename TI>
__device__ __forceinline__ void load_row4(const TI *__restrict__ row, const float *__restrict__ wrow,
                                          int c0, int W, bool vec, float *x) {
  if (vec) {
    if (c0 < W) {
      const pack4<TI> p = *reinterpret_cast<const pack4<TI> *>(row + c0);
#pragma unroll
      for (int k = 0; k < 4; ++k) x[k] = to_f32(p.e[k]);
      if (wrow) {
        const float4 ww = *reinterpret_cast<const float4 *>(wrow + c0);
        x[0] = x[0] * ww.x;
        x[1] = x[1] * ww.y;
        x[2] = x[2] * ww.z;
        x[3] = x[3] * ww.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) x[k] = 0.0f;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = c0 + k;
      float a = 0.0f;
      if (c < W) {
        a = to_f32(row[c]);
        if (wrow) a = a * wrow[c];
      }
      x[k] = a;
    }
  }
}

template <typename TO>
__device__ __forceinline__ void store_row4(TO *__restrict__ row, int c0, int W, bool vec,
                                           const float *y) {
  if (vec) {
    if (c0 < W) {
      pack4<TO> p;
#pragma unroll
      for (int k = 0; k < 4; ++k) store_as(y[k], &p.e[k]);
      *reinterpret_cast<pack4<TO> *>(row + c0) = p;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (c0 + k < W) store_as(y[k], row + c0 + k);
  }
}

// spec layout (floats, 16-byte aligned): [0, L) the spectrum H[k] tau / L in lane order
// (entry 8 j + q: value q of lane j after the forward's last pass); [L, L + 4 E) the E
// twiddle pairs (w, w^2) of the passes (fft_tw_*); [L + 4 E, L + 4 E + W) column weights.
template <int M, typename TI, typename TO>
__global__ void __launch_bounds__((1 << M) / 8)
    ramp_fft_kernel(const TI *__restrict__ in, const float *__restrict__ w,
                    const float *__restrict__ spec, TO *__restrict__ out, int det_h, int W,
                    int pairs, int vec) {
  using F = FftPlan<M>;
  using I = FftInverse<M>;
  constexpr int L = F::L, R0 = F::radix(0);
  __shared__ cpx buf[L + L / kLdsRow];
  const int lane = threadIdx.x;
  const float4 *tw = reinterpret_cast<const float4 *>(spec + L);
  const float *colw = spec + L + 4 * fft_tw_entries(M);
  const int img = (int)(blockIdx.x / (unsigned)pairs);
  const int ra = 2 * (int)(blockIdx.x - (unsigned)img * (unsigned)pairs);
  const bool has_b = ra + 1 < det_h;
  const size_t row_a = ((size_t)img * det_h + ra) * (size_t)W;
  const float *wa = w ? w + (size_t)ra * W : nullptr;
  const int c0 = 4 * lane;  // this lane's four columns; the L / 8 lanes cover [0, L / 2)
  cpx v[8];

  // the weighted rows as one complex sequence (row a real, row b imaginary) in LDS
  {
    float a[4], b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    load_row4(in + row_a, wa, c0, W, vec != 0, a);
    if (has_b) load_row4(in + row_a + W, wa ? wa + W : nullptr, c0, W, vec != 0, b);
#pragma unroll
    for (int k = 0; k < 4; ++k) buf[pad32(c0 + k)] = {a[k], b[k]};
  }
  __syncthreads();
  // forward transform; the first pass reads the first half, the second is the zero padding
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if ((q % R0) * 2 < R0) v[q] = buf[pad32(fft_natural<R0, L>(q, lane))];
  fft_first_pass_half<R0>(v);
  fft_forward_passes<M, 1>(v, buf, lane, tw);
  // the spectrum is in natural order in v: replace X[f] by conj(H[f] X[f]) (H is real), then
  // run the inverse's first pass (same radix, Ns = 1) on the same values
  {
    const float4 *Hl = reinterpret_cast<const float4 *>(spec) + 2 * lane;
    const float4 h0 = Hl[0], h1 = Hl[1];
    const float h[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = {h[q] * v[q].x, -(h[q] * v[q].y)};
  }
  fft_twiddle_dft<I::radix(0), 1, L>(v, lane, tw);
  fft_inverse_passes<M, 1>(v, buf, lane, tw);
  // natural order again (the last pass has radix R0); positions [0, L / 2) are kept, and go
  // through LDS to the lanes that own their columns.  The filtered rows are conj(v).
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if ((q % R0) * 2 < R0) buf[pad32(fft_natural<R0, L>(q, lane))] = v[q];
  __syncthreads();
  float cw[4], ya[4], yb[4];
  if (vec && c0 < W) {
    const float4 c4 = *reinterpret_cast<const float4 *>(colw + c0);
    cw[0] = c4.x;
    cw[1] = c4.y;
    cw[2] = c4.z;
    cw[3] = c4.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) cw[k] = c0 + k < W ? colw[c0 + k] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const cpx o = buf[pad32(c0 + k)];
    ya[k] = o.x * cw[k];
    yb[k] = -o.y * cw[k];
  }
  store_row4(out + row_a, c0, W, vec != 0, ya);
  if (has_b) store_row4(out + row_a + W, c0, W, vec != 0, yb);
}

// One lane per table entry; double precision throughout.
__global__ void ramp_spectrum_kernel(float *__restrict__ spec, int W, int m, double tau,
                                     const float *__restrict__ colw) {
  const int L = 1 << m, P = fft_passes(m), E = fft_tw_entries(m);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const double pi = 3.14159265358979323846;
  if (t < L) {
    // entry 8 j + q: value q of lane j after the forward's last pass (radix RL), whose
    // natural position is k = j + (q / RL) L / 8 + (q % RL) L / RL
    const int RL = fft_radix(m, P - 1), j = t >> 3, q = t & 7;
    const int k = j + (q / RL) * (L / 8) + (q % RL) * (L / RL);
    // H[k] = h(0) + 2 sum_{odd n < W} h(n) cos(2 pi k n / L), h(n) = -1 / (n pi tau)^2
    double acc = 0.0;
    for (int n = 1; n < W; n += 2) {
      const long long kn = ((long long)k * n) & (L - 1);  // the cosine's period is L
      acc += cospi(2.0 * (double)kn / (double)L) / ((double)n * (double)n);
    }
    const double H = 1.0 / (4.0 * tau * tau) - 2.0 * acc / (pi * pi * tau * tau);
    spec[t] = (float)(H * tau / (double)L);
  }
  if (t < E) {
    // entry o of the pass (radix R, Ns) it falls in: w = e^{-2 pi i o / (Ns R)} and w^2,
    // i.e. e^{-2 pi i s / L} for s = o L / (Ns R) and 2 s
    int R = 0, ns = 0, o = t;
    for (int i = 1; i < P && R == 0; ++i) {
      const int n = fft_ns(m, i);
      if (o < n) {
        R = fft_radix(m, i);
        ns = n;
      } else {
        o -= n;
      }
    }
    for (int i = 1; i < P && R == 0; ++i) {
      const int n = ifft_ns(m, i);
      if (o < n) {
        R = ifft_radix(m, i);
        ns = n;
      } else {
        o -= n;
      }
    }
    const long long s = (long long)o * (L / (ns * R));
    double sn, cs;
    sincospi(2.0 * (double)s / (double)L, &sn, &cs);
    spec[L + 4 * t] = (float)cs;
    spec[L + 4 * t + 1] = (float)(-sn);
    sincospi(2.0 * (double)(2 * s) / (double)L, &sn, &cs);
    spec[L + 4 * t + 2] = (float)cs;
    spec[L + 4 * t + 3] = (float)(-sn);
  }
  if (t < W) spec[L + 4 * E + t] = colw ? colw[t] : 1.0f;
}

template <int M, typename TI>
static void launch_ramp_fft_m(const TI *in, const float *w, const float *spec, void *out, int odt,
                              unsigned blocks, int det_h, int W, int pairs, int vec, hipStream_t s) {
  constexpr unsigned threads = 1u << (M - 3);
  if (odt == PYRONN_F32)
    ramp_fft_kernel<M, TI, float><<<blocks, threads, 0, s>>>(in, w, spec, (float *)out, det_h, W, pairs,
                                                            vec);
  else if (odt == PYRONN_BF16)
    ramp_fft_kernel<M, TI, bf16s><<<blocks, threads, 0, s>>>(in, w, spec, (bf16s *)out, det_h, W, pairs,
                                                            vec);
  else
    ramp_fft_kernel<M, TI, _Float16><<<blocks, threads, 0, s>>>(in, w, spec, (_Float16 *)out, det_h, W,
                                                               pairs, vec);
}

static bool aligned_to(const void *p, size_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

template <typename TI>
static void launch_ramp_fft(const TI *in, const float *w, const float *spec, void *out, int odt,
                            int images, int det_h, int W, int m, hipStream_t s) {
  const int pairs = (det_h + 1) / 2;
  const unsigned blocks = (unsigned)((long long)images * pairs);
  // one vector load / store per four samples: every row starts on a multiple of 4 samples
  const size_t out_elem = odt == PYRONN_F32 ? 4 : 2;
  const int vec = W % 4 == 0 && aligned_to(in, 4 * sizeof(TI)) && (!w || aligned_to(w, 16)) &&
                  aligned_to(out, 4 * out_elem);
  switch (m) {
    case 4: launch_ramp_fft_m<4>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 5: launch_ramp_fft_m<5>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 6: launch_ramp_fft_m<6>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 7: launch_ramp_fft_m<7>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 8: launch_ramp_fft_m<8>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 9: launch_ramp_fft_m<9>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 10: launch_ramp_fft_m<10>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 11: launch_ramp_fft_m<11>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    case 12: launch_ramp_fft_m<12>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
    default: launch_ramp_fft_m<13>(in, w, spec, out, odt, blocks, det_h, W, pairs, vec, s); break;
  }
}

}  // namespace pyronn

using namespace pyronn;

extern "C" int pyronn_fdk_ramp_matrix(int det_w, double tau, const float *column_weights,
                                      float *T, pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(det_w > 0 && T, "fdk_ramp_matrix: bad arguments");
  PYRONN_CHECK_ARG(tau > 0.0, "fdk_ramp_matrix: sample spacing must be > 0");
  ramp_matrix_kernel<<<grid_for((long long)det_w * det_w), 256, 0, (hipStream_t)stream>>>(
      T, det_w, tau, column_weights);
  PYRONN_CHECK_LAUNCH("ramp_matrix_kernel");
  return PYRONN_OK;
}

extern "C" int pyronn_fdk_cosine_weights(int det_h, int det_w, double du, double dv, double u0,
                                         double v0, double sdd, float *weights,
                                         pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(det_h > 0 && det_w > 0 && weights, "fdk_cosine_weights: bad arguments");
  PYRONN_CHECK_ARG(sdd > 0.0, "fdk_cosine_weights: SDD must be > 0");
  cosine_weights_kernel<<<grid_for((long long)det_h * det_w), 256, 0, (hipStream_t)stream>>>(
      weights, det_h, det_w, du, dv, u0, v0, sdd);
  PYRONN_CHECK_LAUNCH("cosine_weights_kernel");
  return PYRONN_OK;
}

extern "C" size_t pyronn_fdk_filter_workspace_size(int images, int det_h, int det_w) {
  if (images <= 0 || det_h <= 0 || det_w <= 0) return 0;
  return align_up((size_t)images * det_h * det_w * 4, 256) * 2;
}

extern "C" int pyronn_fdk_filter(const void *sinogram, int sinogram_dtype, int images, int det_h,
                                 int det_w, const float *weights, const float *ramp_matrix,
                                 void *filtered, int filtered_dtype, void *workspace,
                                 size_t workspace_bytes, pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(sinogram && ramp_matrix && filtered && workspace, "fdk_filter: null pointer");
  PYRONN_CHECK_ARG(dtype_ok(sinogram_dtype) && dtype_ok(filtered_dtype), "fdk_filter: bad dtype");
  PYRONN_CHECK_ARG(images > 0 && det_h > 0 && det_w > 0, "fdk_filter: bad shape");
  const long long rows = (long long)images * det_h;
  PYRONN_CHECK_ARG(rows < INT32_MAX, "fdk_filter: too many detector rows for one GEMM");
  const size_t need = pyronn_fdk_filter_workspace_size(images, det_h, det_w);
  if (workspace_bytes < need) {
    set_error("fdk_filter: workspace %zu bytes < required %zu", workspace_bytes, need);
    return PYRONN_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  const long long total = rows * det_w;
  float *weighted = (float *)workspace;
  float *prod = (float *)((char *)workspace + need / 2);
  const long long per_image = (long long)det_h * det_w;
  const dim3 wgrid((unsigned)((per_image + 255) / 256), (unsigned)(images < 512 ? images : 512));
  if (sinogram_dtype == PYRONN_F32)
    weight_kernel<float><<<wgrid, 256, 0, s>>>((const float *)sinogram, weights, weighted, per_image,
                                               images);
  else if (sinogram_dtype == PYRONN_BF16)
    weight_kernel<bf16s><<<wgrid, 256, 0, s>>>((const bf16s *)sinogram, weights, weighted, per_image,
                                               images);
  else
    weight_kernel<_Float16><<<wgrid, 256, 0, s>>>((const _Float16 *)sinogram, weights, weighted,
                                                  per_image, images);
  PYRONN_CHECK_LAUNCH("weight_kernel");

  int herr = 0;
  rocblas_handle h = blas_handle(herr);
  if (herr || !h) {
    set_error("fdk_filter: rocblas handle creation failed");
    return PYRONN_ERR_HIP;
  }
  rocblas_set_stream(h, s);
  const float one = 1.0f, zero = 0.0f;
  float *dst = filtered_dtype == PYRONN_F32 ? (float *)filtered : prod;
  // Row-major out[rows][W] = weighted[rows][W] x T[W][W], i.e. column-major
  // out^T = T^T weighted^T (T is symmetric).
  rocblas_status st = rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, det_w,
                                    (rocblas_int)rows, det_w, &one, ramp_matrix, det_w, weighted,
                                    det_w, &zero, dst, det_w);
  if (st != rocblas_status_success) {
    set_error("fdk_filter: rocblas_sgemm failed (%d)", (int)st);
    return PYRONN_ERR_HIP;
  }
  if (filtered_dtype == PYRONN_BF16)
    convert_kernel<bf16s><<<grid_for(total), 256, 0, s>>>(prod, (bf16s *)filtered, total);
  else if (filtered_dtype == PYRONN_F16)
    convert_kernel<_Float16><<<grid_for(total), 256, 0, s>>>(prod, (_Float16 *)filtered, total);
  PYRONN_CHECK_LAUNCH("convert_kernel");
  return PYRONN_OK;
}

extern "C" size_t pyronn_fdk_ramp_spectrum_size(int det_w) {
  if (det_w <= 0 || det_w > (1 << (kFftMaxLog2 - 1))) return 0;
  const int m = fft_log2(det_w);
  return ((size_t)1 << m) + (size_t)4 * fft_tw_entries(m) + (size_t)det_w;
}

extern "C" int pyronn_fdk_ramp_spectrum(int det_w, double tau, const float *column_weights,
                                        float *spectrum, pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(spectrum, "fdk_ramp_spectrum: null pointer");
  PYRONN_CHECK_ARG(pyronn_fdk_ramp_spectrum_size(det_w) > 0,
                   "fdk_ramp_spectrum: det_w %d outside 1..%d", det_w, 1 << (kFftMaxLog2 - 1));
  PYRONN_CHECK_ARG(tau > 0.0, "fdk_ramp_spectrum: sample spacing must be > 0");
  PYRONN_CHECK_ARG(((uintptr_t)spectrum & 15) == 0, "fdk_ramp_spectrum: table must be 16-byte aligned");
  const int m = fft_log2(det_w);
  int n = (1 << m) > det_w ? (1 << m) : det_w;
  if (fft_tw_entries(m) > n) n = fft_tw_entries(m);
  ramp_spectrum_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(spectrum, det_w, m, tau,
                                                                         column_weights);
  PYRONN_CHECK_LAUNCH("ramp_spectrum_kernel");
  return PYRONN_OK;
}

extern "C" int pyronn_fdk_filter_fft(const void *sinogram, int sinogram_dtype, int images,
                                     int det_h, int det_w, const float *weights,
                                     const float *spectrum, void *filtered, int filtered_dtype,
                                     pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(sinogram && spectrum && filtered, "fdk_filter_fft: null pointer");
  PYRONN_CHECK_ARG(dtype_ok(sinogram_dtype) && dtype_ok(filtered_dtype), "fdk_filter_fft: bad dtype");
  PYRONN_CHECK_ARG(images > 0 && det_h > 0 && det_w > 0, "fdk_filter_fft: bad shape");
  PYRONN_CHECK_ARG(pyronn_fdk_ramp_spectrum_size(det_w) > 0,
                   "fdk_filter_fft: det_w %d outside 1..%d", det_w, 1 << (kFftMaxLog2 - 1));
  PYRONN_CHECK_ARG((long long)images * ((det_h + 1) / 2) <= INT32_MAX,
                   "fdk_filter_fft: too many row pairs for one launch");
  PYRONN_CHECK_ARG(((uintptr_t)spectrum & 15) == 0, "fdk_filter_fft: table must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int m = fft_log2(det_w);
  if (sinogram_dtype == PYRONN_F32)
    launch_ramp_fft((const float *)sinogram, weights, spectrum, filtered, filtered_dtype, images,
                    det_h, det_w, m, s);
  else if (sinogram_dtype == PYRONN_BF16)
    launch_ramp_fft((const bf16s *)sinogram, weights, spectrum, filtered, filtered_dtype, images,
                    det_h, det_w, m, s);
  else
    launch_ramp_fft((const _Float16 *)sinogram, weights, spectrum, filtered, filtered_dtype, images,
                    det_h, det_w, m, s);
  PYRONN_CHECK_LAUNCH("ramp_fft_kernel");
  return PYRONN_OK;
}
