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
