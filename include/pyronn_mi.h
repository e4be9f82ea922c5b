/*
 * pyronn_mi.h - C ABI of the MI355X (gfx950) projector / backprojector library
 * libpyronn_mi.so.
 *
 * One entry point per reference operator.  Each replaces the reference's TF
 * op Compute() plus its CUDA launcher for that operator; the citation above
 * each declaration names the REGISTER_OP input list and the launcher it
 * replaces (paths relative to the reference repository root).
 *
 * Conventions shared by every entry point
 *  - All array arguments are DEVICE pointers, batch-major, densely packed,
 *    exactly as the reference op receives its tensors:
 *      volume_origin / volume_spacing   float[B][3] (z,y,x) for 3D, float[B][2] (y,x) for 2D
 *      projection_matrices              float[B][n][3][4] row-major
 *      ray_vectors / central_ray_vectors float[B][n][2] (x,y)
 *      step_size, projection_multiplier, detector_origin, detector_spacing,
 *      source_2_isocenter_distance, source_2_detector_distance    float[B][1]
 *    Scalars stay on the device (the reference dereferences them in-kernel);
 *    shapes are host integers (the reference copies its shape tensor to the
 *    host).  Outputs are always float32 and are overwritten, never accumulated.
 *  - `*_dtype` selects the storage type of the data input: PYRONN_F32,
 *    PYRONN_BF16 or PYRONN_F16.  Arithmetic is float32 in every case.
 *  - Calls are asynchronous on `stream` (NULL = the null stream), allocate no
 *    device memory and keep no state between calls: the caller passes the
 *    workspace sized by the matching *_workspace_size() function, 16-byte
 *    aligned (every device allocation is).  Calls are reentrant; two calls may
 *    run concurrently if their workspaces differ.
 *  - Return value: PYRONN_OK or an error code; pyronn_last_error() gives the
 *    message for the calling thread.  Invalid arguments are rejected before
 *    anything is launched.  Invalid *values* that only the device sees (a
 *    singular projection matrix, a step size <= 0 or NaN) produce NaN in the
 *    affected output elements instead of a hang or a process exit.
 */
#ifndef PYRONN_MI_H
#define PYRONN_MI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *pyronn_stream_t; /* == hipStream_t */

enum pyronn_status {
  PYRONN_OK = 0,
  PYRONN_ERR_INVALID_ARGUMENT = 1,
  PYRONN_ERR_WORKSPACE = 2,
  PYRONN_ERR_TOO_LARGE = 3,
  PYRONN_ERR_HIP = 4
};

enum pyronn_dtype { PYRONN_F32 = 0, PYRONN_BF16 = 1, PYRONN_F16 = 2 };

/* Largest number of problems in one batched (persistent) launch. */
#define PYRONN_MAX_BATCHED_PROBLEMS 16

/* Message for the last failing call on this thread ("" if none). */
const char *pyronn_last_error(void);

/* Library version as 10000*major + 100*minor + patch. */
int pyronn_version(void);

/*
 * Per-view cone geometry on the device, in FP64, from the projection matrices:
 * inv_ar[b][v] = diag(1/spacing_xyz) * M^-1 (row-major 3x3) and
 * src[b][v] = (C - origin_xyz) / spacing_xyz with C = -M^-1 p4 the camera
 * centre.  Singular M gives NaN rows.
 * Replaces ConeProjection3DOp::DecomposeProjectionMatrices
 * (cc/ops/cone_projector_3D_OPKernel.cc:96-182) and Geometry::getCameraCenter
 * (cc/helper_headers/helper_geometry_cpu.h:30-43), which ran on the host with
 * Eigen's JacobiSVD after synchronous device-to-host copies.
 */
int pyronn_cone_geometry_setup(const float *projection_matrices, const float *volume_origin,
                               const float *volume_spacing, int batch, int n_views, float *inv_ar,
                               float *src, pyronn_stream_t stream);

/* Workspace bytes for pyronn_cone_projection3d with these arguments. */
size_t pyronn_cone_projection3d_workspace_size(int batch, int D, int H, int W, int n_views,
                                               int volume_dtype);

/*
 * ConeProjection3D: volume [B][D][H][W] -> sinogram [B][n][H_d][W_d] (float32).
 * Ray-driven trilinear line integrals, step `step_size` in voxel units,
 * hardware_interp selects the texture-mode clip box [-0.5, N-0.5] (exact
 * weights) instead of [0, N].  projection_multiplier is accepted and ignored,
 * as in the reference.
 * Replaces REGISTER_OP("ConeProjection3D") (cc/ops/cone_projector_3D_OPKernel.cc:34-60),
 * ConeProjection3DOp::Compute (:185-248), Cone_Projection_Kernel_Launcher
 * (cc/kernels/cone_projector_3D_CudaKernel.cu.cc:225-257) and
 * Cone_Projection_Kernel_Tex_Interp_Launcher
 * (cc/kernels/cone_projector_3D_CudaKernel_hardware_interp.cu.cc:199-262).
 */
int pyronn_cone_projection3d(const void *volume, int volume_dtype, int batch, int D, int H, int W,
                             const float *volume_origin, const float *volume_spacing,
                             const float *projection_matrices, const float *step_size,
                             const float *projection_multiplier, int n_views, int det_h, int det_w,
                             int hardware_interp, float *sinogram, void *workspace,
                             size_t workspace_bytes, pyronn_stream_t stream);

/*
 * pyronn_cone_projection3d without a workspace: it allocates no device memory and takes
 * none from the caller.  No counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results bit for bit those of pyronn_cone_projection3d.
 * Its only scratch is the per-view geometry, held in a fixed device array of 3 MiB per
 * GPU (kept until the process ends); the volume is sampled in place rather than through a
 * cell copy, so it runs slower than pyronn_cone_projection3d.
 * Limit: batch * n_views <= 65536, else PYRONN_ERR_TOO_LARGE.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (on different streams, or from
 * different threads) race and must be serialised by the caller.  The stream must belong
 * to the current device.
 */
int pyronn_cone_projection3d_noalloc(const void *volume, int volume_dtype, int batch, int D, int H,
                                     int W, const float *volume_origin, const float *volume_spacing,
                                     const float *projection_matrices, const float *step_size,
                                     const float *projection_multiplier, int n_views, int det_h,
                                     int det_w, int hardware_interp, float *sinogram,
                                     pyronn_stream_t stream);

/* Workspace bytes for pyronn_cone_backprojection3d with these arguments. */
size_t pyronn_cone_backprojection3d_workspace_size(int batch, int n_views, int det_h, int det_w,
                                                   int sinogram_dtype);

/*
 * ConeBackprojection3D: sinogram [B][n][H_d][W_d] -> volume [B][D][H][W] (float32).
 * Voxel-driven: vol = projection_multiplier * sum_n bilinear(sino_n, P_n x) / w_n^2,
 * zero outside the detector (the reference's off-by-one bounds test, Q11, is
 * not reproduced).  step_size is accepted and ignored, as in the reference.
 * Replaces REGISTER_OP("ConeBackprojection3D") (cc/ops/cone_backprojector_3D_OPKernel.cc:27-53),
 * ConeBackprojection3DOp::Compute (:75-140), Cone_Backprojection3D_Kernel_Launcher
 * (cc/kernels/cone_backprojector_3D_CudaKernel.cu.cc:121-146) and
 * Cone_Backprojection3D_Kernel_Tex_Interp_Launcher
 * (cc/kernels/cone_backprojector_3D_CudaKernel_hardware_interp.cu.cc:104-160).
 */
int pyronn_cone_backprojection3d(const void *sinogram, int sinogram_dtype, int batch, int n_views,
                                 int det_h, int det_w, const float *volume_origin,
                                 const float *volume_spacing, const float *projection_matrices,
                                 const float *step_size, const float *projection_multiplier, int D,
                                 int H, int W, int hardware_interp, float *volume, void *workspace,
                                 size_t workspace_bytes, pyronn_stream_t stream);

/*
 * Backprojection of one range of views into a running sum.  No counterpart in the
 * reference: it serves callers that receive a sinogram in view chunks (a server that
 * overlaps the upload of chunk k + 1 with the backprojection of chunk k).  sinogram and
 * projection_matrices hold the chunk's n_views views; the workspace is sized by
 * pyronn_cone_backprojection3d_workspace_size for the chunk.  mode bit 0: continue the
 * per-voxel running sums stored in volume by the previous chunk (else start from zero);
 * bit 1: multiply by projection_multiplier when done (the last chunk).  Chunks of one
 * sinogram processed in view order with modes 0, 1, ..., 1, 3 give bit for bit the
 * result of pyronn_cone_backprojection3d over all views (mode 2 alone is that operator):
 * every voxel accumulates its views in the same order and the running sum between chunks
 * is the same float the one-pass accumulation holds.
 * Only the z slices [z_begin, z_end) of the D-slice volume are computed and written (each
 * voxel's value does not depend on the range), so a caller can split the last chunk
 * into z slabs and download each slab while the next one is computed.  mode bit 2: the
 * workspace still holds this chunk's detector cells from the previous call (same
 * sinogram, same workspace, e.g. the previous z slab); skips rebuilding them.
 */
int pyronn_cone_backprojection3d_views(const void *sinogram, int sinogram_dtype, int batch,
                                       int n_views, int det_h, int det_w,
                                       const float *volume_origin, const float *volume_spacing,
                                       const float *projection_matrices,
                                       const float *projection_multiplier, int D, int H, int W,
                                       int hardware_interp, int mode, int z_begin, int z_end,
                                       float *volume,
                                       void *workspace, size_t workspace_bytes,
                                       pyronn_stream_t stream);

/*
 * ParallelProjection2D: volume [B][H][W] -> sinogram [B][n][det].
 * step_factor scales min(spacing) to the march step; the reference hard-codes
 * 0.2 (cc/kernels/par_projector_2D_CudaKernel.cu.cc:166), pass 0.2 for parity.
 * Replaces REGISTER_OP("ParallelProjection2D") (cc/ops/par_projector_2D_OPKernel.cc:30-56),
 * ParallelProjection2DOp::Compute (:67-120) and Parallel_Projection2D_Kernel_Launcher
 * (cc/kernels/par_projector_2D_CudaKernel.cu.cc:149-178).
 */
int pyronn_parallel_projection2d(const void *volume, int volume_dtype, int batch, int H, int W,
                                 const float *volume_origin, const float *detector_origin,
                                 const float *volume_spacing, const float *detector_spacing,
                                 const float *ray_vectors, int n_views, int det,
                                 float step_factor, float *sinogram, pyronn_stream_t stream);

/*
 * ParallelBackprojection2D: sinogram [B][n][det] -> volume [B][H][W].
 * Replaces REGISTER_OP("ParallelBackprojection2D") (cc/ops/par_backprojector_2D_OPKernel.cc:26-51),
 * its Compute (:63-115) and Parallel_Backprojection2D_Kernel_Launcher
 * (cc/kernels/par_backprojector_2D_CudaKernel.cu.cc:77-105).
 */
int pyronn_parallel_backprojection2d(const void *sinogram, int sinogram_dtype, int batch,
                                     int n_views, int det, const float *volume_origin,
                                     const float *detector_origin, const float *volume_spacing,
                                     const float *detector_spacing, const float *ray_vectors,
                                     int H, int W, float *volume, pyronn_stream_t stream);

/*
 * FanProjection2D: volume [B][H][W] -> sinogram [B][n][det].
 * Replaces REGISTER_OP("FanProjection2D") (cc/ops/fan_projector_2D_OPKernel.cc:26-53),
 * FanProjection2DOp::Compute (:66-128) and Fan_Projection_Kernel_Launcher
 * (cc/kernels/fan_projector_2D_CudaKernel.cu.cc:165-194).
 */
int pyronn_fan_projection2d(const void *volume, int volume_dtype, int batch, int H, int W,
                            const float *volume_origin, const float *detector_origin,
                            const float *volume_spacing, const float *detector_spacing,
                            const float *source_2_isocenter_distance,
                            const float *source_2_detector_distance,
                            const float *central_ray_vectors, int n_views, int det,
                            float step_factor, float *sinogram, pyronn_stream_t stream);

/*
 * FanBackprojection2D: sinogram [B][n][det] -> volume [B][H][W].
 * Replaces REGISTER_OP("FanBackprojection2D") (cc/ops/fan_backprojector_2D_OPKernel.cc:26-53),
 * its Compute (:67-128) and Fan_Backprojection2D_Kernel_Launcher
 * (cc/kernels/fan_backprojector_2D_CudaKernel.cu.cc:83-113).
 */
int pyronn_fan_backprojection2d(const void *sinogram, int sinogram_dtype, int batch, int n_views,
                                int det, const float *volume_origin, const float *detector_origin,
                                const float *volume_spacing, const float *detector_spacing,
                                const float *source_2_isocenter_distance,
                                const float *source_2_detector_distance,
                                const float *central_ray_vectors, int H, int W, float *volume,
                                pyronn_stream_t stream);

/* --------------------------------------------------------------------------
 * Batched cone operators over problems of different shapes and geometries.
 * No single counterpart in the reference: ConeProjection3DOp::Compute
 * (cc/ops/cone_projector_3D_OPKernel.cc:234-247) loops over batch items of one
 * shape and launches once per item.  Here up to PYRONN_MAX_BATCHED_PROBLEMS
 * volumes / sinograms of any shapes run in ONE launch (plus small per-problem
 * layout passes; a forward of more than 2^24 - 1 workgroups, one per item, takes
 * several launches in item order, and such a backprojection the persistent grid);
 * every output equals the single-problem call bit for bit.  All problems share
 * the storage dtype and hardware_interp.
 * -------------------------------------------------------------------------- */
typedef struct pyronn_cone_problem {
  const void *input;                  /* forward: volume [D][H][W]; backprojection: sinogram [n][H_d][W_d] */
  float *output;                      /* forward: sinogram [n][H_d][W_d]; backprojection: volume [D][H][W] */
  const float *volume_origin;         /* [3] (z,y,x) */
  const float *volume_spacing;        /* [3] (z,y,x) */
  const float *projection_matrices;   /* [n][3][4] */
  const float *step_size;             /* [1]  (read by the forward) */
  const float *projection_multiplier; /* [1]  (read by the backprojection) */
  int D, H, W, n_views, det_h, det_w;
} pyronn_cone_problem;

size_t pyronn_cone_projection3d_batched_workspace_size(const pyronn_cone_problem *problems, int count,
                                                       int dtype);
/* Tuning parameters (results are the same bits for every value):
 * blocks_per_cu > 0: a persistent grid of that many workgroups per CU claiming items from
 *   per-XCD counters; <= 0 (default): one workgroup per item in item order (no claims);
 * forward_tile: the detector tile of a 256-lane workgroup and the waves within it,
 *   0: 16 x 16 pixels, waves of 4 columns x 16 rows (the single-problem shape of
 *      16-bit volumes above 2^25 voxels),
 *   1: 8 x 32, waves 2 x 32;
 *   2: 32 x 8, waves 8 x 8 (the single-problem shape of f32 and smaller 16-bit volumes);
 *   3: 16 x 16, waves 8 x 8. */
int pyronn_cone_projection3d_batched(const pyronn_cone_problem *problems, int count, int dtype,
                                     int hardware_interp, int blocks_per_cu, int forward_tile,
                                     void *workspace, size_t workspace_bytes, pyronn_stream_t stream);

size_t pyronn_cone_backprojection3d_batched_workspace_size(const pyronn_cone_problem *problems,
                                                           int count, int dtype);
/* z_per_lane: voxels along z per lane, 4 or 8 (<= 0: 8 for f32 and bf16, 4 for f16 storage).
 * blocks_per_cu > 0: a persistent grid of that many workgroups per CU claiming items from
 * per-XCD counters; <= 0 (default): one workgroup per item in item order (the persistent
 * grid of 8 per CU when the items exceed one dispatch's 2^24 - 1 workgroups).  Both give the
 * same bits; one workgroup per item keeps the single kernel's register budget and was
 * measured faster (200.6 against 217.2 ms on an eight-problem f32 batch). */
int pyronn_cone_backprojection3d_batched(const pyronn_cone_problem *problems, int count, int dtype,
                                         int hardware_interp, int z_per_lane, int blocks_per_cu,
                                         void *workspace, size_t workspace_bytes,
                                         pyronn_stream_t stream);

/* --------------------------------------------------------------------------
 * FDK / FBP pre-processing.  No counterpart in the reference: its FDK weights
 * and filters live in the companion pyronn package (not part of
 * PYRO-NN-Layers); the backprojectors above expect projections already
 * weighted and filtered, with the FDK normalisation passed as
 * projection_multiplier (cc/kernels/cone_backprojector_3D_CudaKernel.cu.cc:117).
 * Formulas: Kak & Slaney, ch. 3 (Ram-Lak kernel, cosine pre-weighting).
 * -------------------------------------------------------------------------- */

/* T[j][c] = tau * h(c - j) * column_weights[c]: the zero-padded Ram-Lak filter
 * of a row of det_w samples spaced tau (mm) as a Toeplitz matrix (device,
 * det_w^2 floats), with an optional per-output-sample weight folded in
 * (column_weights: device float[det_w] or NULL). */
int pyronn_fdk_ramp_matrix(int det_w, double tau, const float *column_weights, float *T,
                           pyronn_stream_t stream);

/* weights[r][c] = sdd / sqrt(sdd^2 + ((c - u0) du)^2 + ((r - v0) dv)^2)
 * (device, det_h * det_w floats; det_h = 1, dv = 0 for a fan-beam row). */
int pyronn_fdk_cosine_weights(int det_h, int det_w, double du, double dv, double u0, double v0,
                              double sdd, float *weights, pyronn_stream_t stream);

/* Workspace bytes for pyronn_fdk_filter. */
size_t pyronn_fdk_filter_workspace_size(int images, int det_h, int det_w);

/* filtered[i][r][:] = (weights[r][:] .* sinogram[i][r][:]) x T for each of the
 * `images` projection images (weights may be NULL: no weighting).  The row
 * product runs as one rocBLAS SGEMM (a rocBLAS handle is created on the first
 * call per thread and device); the output is stored as filtered_dtype. */
int pyronn_fdk_filter(const void *sinogram, int sinogram_dtype, int images, int det_h, int det_w,
                      const float *weights, const float *ramp_matrix, void *filtered,
                      int filtered_dtype, void *workspace, size_t workspace_bytes,
                      pyronn_stream_t stream);

/* The same filter as pyronn_fdk_filter, evaluated by FFT in one fused kernel (no
 * counterpart in the reference, whose filtering lived outside the op library).
 * A row of det_w samples is convolved circularly at length L, the smallest power of
 * two >= max(16, 2 det_w - 1), so the zero padding makes it the linear convolution
 * with h(c - j), |c - j| < det_w, that T holds.  det_w is limited to 1..4096.
 *
 * pyronn_fdk_ramp_spectrum_size: floats of the filter table for det_w (0 outside
 * the limit).  pyronn_fdk_ramp_spectrum: builds the table (device, 16-byte aligned):
 * the ramp spectrum scaled by tau / L, the FFT twiddles and the per-output-sample
 * column weights (NULL: all 1), the FFT counterpart of pyronn_fdk_ramp_matrix.  The
 * table's layout is private to the library; a table is valid for the det_w it was
 * built for.
 *
 * pyronn_fdk_filter_fft: filtered[i][r][:] = ((weights[r][:] .* sinogram[i][r][:])
 * conv h) .* column_weights, stored as filtered_dtype; no workspace.  Rows 2p and
 * 2p + 1 of an image are filtered together (as one complex sequence), so a row's
 * result depends only on its own image, whatever the number of images per call; a NaN or
 * infinity in one of the two rows makes both filtered rows non-finite (the GEMM filter:
 * only its own row).  Results agree with the GEMM filter to float rounding (about 1e-7
 * relative). */
size_t pyronn_fdk_ramp_spectrum_size(int det_w);
int pyronn_fdk_ramp_spectrum(int det_w, double tau, const float *column_weights, float *spectrum,
                             pyronn_stream_t stream);
int pyronn_fdk_filter_fft(const void *sinogram, int sinogram_dtype, int images, int det_h,
                          int det_w, const float *weights, const float *spectrum, void *filtered,
                          int filtered_dtype, pyronn_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PYRONN_MI_H */
