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
// 1 / L and tau), conjugated at the end.  The forward transform's first pass reads the
// weighted input straight from global memory; its last pass leaves the spectrum in
// natural order in exactly the registers that the inverse transform's first pass (same
// radix, no twiddles) reads, so the spectral product needs no LDS round trip; the
// inverse's last pass writes natural order straight to global memory.  The kernel is
// instantiated per transform size, so every radix, stride and LDS offset is a compile-time
// constant (one address computation per butterfly instead of one per value: measured
// 0.39 -> 0.27 ms per 90-view 1024 x 768 range).
constexpr int kFftMinLog2 = 4, kFftMaxLog2 = 13;  // 16 .. 8192 points, 2 .. 1024 lanes

static int fft_log2(int det_w) {
  int m = kFftMinLog2;
  while ((1LL << m) < 2LL * det_w - 1) ++m;
  return m;
}

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
__device__ __forceinline__ void dft8(cpx *v) {
  cpx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cpx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
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

// Transform of L = 2^M points: one radix-2 / radix-4 pass first when M is not a multiple
// of 3, radix 8 after; Ns(i) = product of the radices before pass i.
template <int M>
struct FftPlan {
  static constexpr int L = 1 << M, T = L >> 3, passes = (M + 2) / 3;
  static constexpr int radix(int i) { return (M % 3 == 0 || i > 0) ? 8 : (M % 3 == 1 ? 2 : 4); }
  static constexpr int ns(int i) {
    int n = 1;
    for (int k = 0; k < i; ++k) n *= radix(k);
    return n;
  }
};

// LDS slot (one complex value, 8 bytes) of sequence position p: one pad slot per 32, which
// spreads the strided stores of the early passes over the banks.
__device__ __forceinline__ int pad32(int p) { return p + (p >> 5); }

// Butterfly b (< 8 / R) of a lane in a pass of radix R has index j = lane + b L / 8; its
// values v[b R + r] come from x[j + r L / R] and go to x[(j / Ns) Ns R + (j % Ns) + r Ns].
template <int R, int NS, int L>
__device__ __forceinline__ void fft_twiddle_dft(cpx *v, int lane, const cpx *__restrict__ tw) {
  constexpr int T = L / 8;
  if constexpr (NS > 1) {
#pragma unroll
    for (int b = 0; b < 8 / R; ++b) {
      // w^r, w = e^{-2 pi i k / (Ns R)}: w and w^2 from the table, higher powers as
      // products of at most two table values or one earlier product
      const int k = (lane + b * T) & (NS - 1);
      const int step = k * (L / (NS * R));
      const cpx w1 = tw[step];
      v[b * R + 1] = cmul(v[b * R + 1], w1);
      if constexpr (R >= 4) {
        const cpx w2 = tw[2 * step], w3 = cmul(w2, w1);
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
  if constexpr (NS % 32 == 0) return pad32(base) + r * (NS + NS / 32);
  else if constexpr (R * NS <= 32) return pad32(base) + r * NS;
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
      const int a = (T % 32 == 0 && S % 32 == 0) ? pad32(lane) + (b * T + r * S) / 32 * 33
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
                                                   const cpx *__restrict__ tw) {
  using F = FftPlan<M>;
  if constexpr (I < F::passes) {
    if constexpr (I > 1) __syncthreads();
    fft_store<F::radix(I - 1), F::ns(I - 1), F::L>(v, buf, lane);
    __syncthreads();
    fft_load<F::radix(I), F::L>(v, buf, lane);
    fft_twiddle_dft<F::radix(I), F::ns(I), F::L>(v, lane, tw);
    fft_forward_passes<M, I + 1>(v, buf, lane, tw);
  }
}

// Inverse transform: the forward passes in reverse radix order; pass K of the inverse has
// radix radix(P - 1 - K) and Ns the product of the inverse radices before it.
template <int M>
struct FftInverse {
  using F = FftPlan<M>;
  static constexpr int radix(int k) { return F::radix(F::passes - 1 - k); }
  static constexpr int ns(int k) {
    int n = 1;
    for (int i = 0; i < k; ++i) n *= radix(i);
    return n;
  }
};

template <int M, int K>
__device__ __forceinline__ void fft_inverse_passes(cpx *v, cpx *buf, int lane,
                                                   const cpx *__restrict__ tw) {
  using F = FftPlan<M>;
  using I = FftInverse<M>;
  if constexpr (K < F::passes) {
    __syncthreads();
    fft_store<I::radix(K - 1), I::ns(K - 1), F::L>(v, buf, lane);
    __syncthreads();
    fft_load<I::radix(K), F::L>(v, buf, lane);
    fft_twiddle_dft<I::radix(K), I::ns(K), F::L>(v, lane, tw);
    fft_inverse_passes<M, K + 1>(v, buf, lane, tw);
  }
}

// spec layout (floats): [0, L) H[k] tau / L; [L, 3L) twiddles e^{-2 pi i m / L} as
// (re, im) pairs; [3L, 3L + W) column weights.
template <int M, typename TI, typename TO>
__global__ void __launch_bounds__((1 << M) / 8)
    ramp_fft_kernel(const TI *__restrict__ in, const float *__restrict__ w,
                    const float *__restrict__ spec, TO *__restrict__ out, int det_h, int W,
                    int pairs) {
  using F = FftPlan<M>;
  using I = FftInverse<M>;
  constexpr int L = F::L, R0 = F::radix(0), RL = F::radix(F::passes - 1);
  __shared__ cpx buf[L + L / 32];
  const int lane = threadIdx.x;
  const float *H = spec;
  const cpx *tw = reinterpret_cast<const cpx *>(spec + L);
  const float *colw = spec + 3 * L;
  const int img = (int)(blockIdx.x / (unsigned)pairs);
  const int ra = 2 * (int)(blockIdx.x - (unsigned)img * (unsigned)pairs);
  const bool has_b = ra + 1 < det_h;
  const size_t row_a = ((size_t)img * det_h + ra) * (size_t)W;
  const float *wa = w ? w + (size_t)ra * W : nullptr;
  cpx v[8];

  // forward transform; its first pass reads the weighted, zero-padded rows from global memory
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int c = fft_natural<R0, L>(q, lane);
    float a = 0.0f, b = 0.0f;
    if (c < W) {
      a = to_f32(in[row_a + c]);
      if (wa) a = a * wa[c];
      if (has_b) {
        b = to_f32(in[row_a + W + c]);
        if (wa) b = b * wa[W + c];
      }
    }
    v[q] = {a, b};
  }
  fft_twiddle_dft<R0, 1, L>(v, lane, tw);
  fft_forward_passes<M, 1>(v, buf, lane, tw);
  // the spectrum is in natural order in v: replace X[f] by conj(H[f] X[f]) (H is real), then
  // run the inverse's first pass (same radix, Ns = 1) on the same values
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const float h = H[fft_natural<RL, L>(q, lane)];
    v[q] = {h * v[q].x, -(h * v[q].y)};
  }
  fft_twiddle_dft<I::radix(0), 1, L>(v, lane, tw);
  fft_inverse_passes<M, 1>(v, buf, lane, tw);
  // natural order again; the filtered rows are conj(v)
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int c = fft_natural<I::radix(F::passes - 1), L>(q, lane);
    if (c < W) {
      store_as(v[q].x * colw[c], out + row_a + c);
      if (has_b) store_as(-v[q].y * colw[c], out + row_a + W + c);
    }
  }
}

// One lane per table entry; double precision throughout.
__global__ void ramp_spectrum_kernel(float *__restrict__ spec, int W, int m, double tau,
                                     const float *__restrict__ colw) {
  const int L = 1 << m;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const double pi = 3.14159265358979323846;
  if (t < L) {
    // H[k] = h(0) + 2 sum_{odd n < W} h(n) cos(2 pi k n / L), h(n) = -1 / (n pi tau)^2
    double acc = 0.0;
    for (int n = 1; n < W; n += 2) {
      const long long kn = ((long long)t * n) & (L - 1);  // the cosine's period is L
      acc += cospi(2.0 * (double)kn / (double)L) / ((double)n * (double)n);
    }
    const double H = 1.0 / (4.0 * tau * tau) - 2.0 * acc / (pi * pi * tau * tau);
    spec[t] = (float)(H * tau / (double)L);
    double sn, cs;
    sincospi(2.0 * (double)t / (double)L, &sn, &cs);
    spec[L + 2 * t] = (float)cs;
    spec[L + 2 * t + 1] = (float)(-sn);
  }
  if (t < W) spec[3 * L + t] = colw ? colw[t] : 1.0f;
}

template <int M, typename TI>
static void launch_ramp_fft_m(const TI *in, const float *w, const float *spec, void *out, int odt,
                              unsigned blocks, int det_h, int W, int pairs, hipStream_t s) {
  constexpr unsigned threads = 1u << (M - 3);
  if (odt == PYRONN_F32)
    ramp_fft_kernel<M, TI, float><<<blocks, threads, 0, s>>>(in, w, spec, (float *)out, det_h, W, pairs);
  else if (odt == PYRONN_BF16)
    ramp_fft_kernel<M, TI, bf16s><<<blocks, threads, 0, s>>>(in, w, spec, (bf16s *)out, det_h, W, pairs);
  else
    ramp_fft_kernel<M, TI, _Float16><<<blocks, threads, 0, s>>>(in, w, spec, (_Float16 *)out, det_h, W,
                                                               pairs);
}

template <typename TI>
static void launch_ramp_fft(const TI *in, const float *w, const float *spec, void *out, int odt,
                            int images, int det_h, int W, int m, hipStream_t s) {
  const int pairs = (det_h + 1) / 2;
  const unsigned blocks = (unsigned)((long long)images * pairs);
  switch (m) {
    case 4: launch_ramp_fft_m<4>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 5: launch_ramp_fft_m<5>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 6: launch_ramp_fft_m<6>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 7: launch_ramp_fft_m<7>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 8: launch_ramp_fft_m<8>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 9: launch_ramp_fft_m<9>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 10: launch_ramp_fft_m<10>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 11: launch_ramp_fft_m<11>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    case 12: launch_ramp_fft_m<12>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
    default: launch_ramp_fft_m<13>(in, w, spec, out, odt, blocks, det_h, W, pairs, s); break;
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
  return (size_t)3 * ((size_t)1 << fft_log2(det_w)) + (size_t)det_w;
}

extern "C" int pyronn_fdk_ramp_spectrum(int det_w, double tau, const float *column_weights,
                                        float *spectrum, pyronn_stream_t stream) {
  clear_error();
  PYRONN_CHECK_ARG(spectrum, "fdk_ramp_spectrum: null pointer");
  PYRONN_CHECK_ARG(pyronn_fdk_ramp_spectrum_size(det_w) > 0,
                   "fdk_ramp_spectrum: det_w %d outside 1..%d", det_w, 1 << (kFftMaxLog2 - 1));
  PYRONN_CHECK_ARG(tau > 0.0, "fdk_ramp_spectrum: sample spacing must be > 0");
  PYRONN_CHECK_ARG(((uintptr_t)spectrum & 7) == 0, "fdk_ramp_spectrum: table must be 8-byte aligned");
  const int m = fft_log2(det_w);
  const int n = (1 << m) > det_w ? (1 << m) : det_w;
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
  PYRONN_CHECK_ARG(((uintptr_t)spectrum & 7) == 0, "fdk_filter_fft: table must be 8-byte aligned");
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
