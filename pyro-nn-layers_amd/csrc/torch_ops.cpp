// Torch op library: the six reference operators registered under the
// `pyronn_layers` namespace with the reference's input lists, shape rule
// (output = input.shape[0:1] ++ shape argument) and attributes.  Each op
// validates and packs its tensors, then calls the C ABI of libpyronn_mi.so on
// torch's current HIP stream.  Only a CUDA (HIP) kernel is registered, as the
// reference registers only DEVICE_GPU kernels (e.g.
// cc/ops/par_projector_2D_OPKernel.cc:123); a CPU tensor is an error.
//
// Differences from the reference op layer (cc/ops/*_OPKernel.cc):
//  - shape arguments are int lists (the reference copies an int32 shape tensor
//    to the host; the Python wrappers accept either form);
//  - geometry tensors may omit the batch dimension and are then shared by
//    every batch item;
//  - the whole batch is one launch (the reference loops over batch items);
//  - the volume / sinogram may be float32, bfloat16 or float16; outputs are
//    float32;
//  - the 2D projectors take an optional step_size (x min spacing, default 0.2,
//    the reference's hard-coded value).
#include <ATen/ATen.h>
#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <vector>

#include "pyronn_mi.h"

namespace {

int dtype_code(const at::Tensor &t, const char *op) {
  switch (t.scalar_type()) {
    case at::kFloat: return PYRONN_F32;
    case at::kBFloat16: return PYRONN_BF16;
    case at::kHalf: return PYRONN_F16;
    default: TORCH_CHECK(false, op, ": data tensor must be float32, bfloat16 or float16, got ",
                         t.scalar_type());
  }
  return -1;
}

// Geometry tensor as float32 [B, shape...] contiguous on `dev`.  Accepts the
// per-item shape without the batch dimension (broadcast to every item).
at::Tensor geom(const at::Tensor &t, int64_t B, std::vector<int64_t> shape, const char *op,
                const char *name, const at::Device &dev) {
  int64_t per = 1;
  for (auto s : shape) per *= s;
  at::Tensor g = t.to(dev, at::kFloat).contiguous();
  if (g.numel() == per) {
    g = g.reshape({1, per}).expand({B, per}).contiguous();
  }
  TORCH_CHECK(g.numel() == B * per, op, ": ", name, " has ", g.numel(), " elements, expected ", B,
              " x ", per);
  std::vector<int64_t> full{B};
  full.insert(full.end(), shape.begin(), shape.end());
  return g.reshape(full);
}

// An empty output (no batch items, no views, a zero-sized volume or detector) needs no
// launch; the reference op also returns an empty tensor for an empty batch.
bool empty_output(const at::Tensor &out) { return out.numel() == 0; }

void check_rc(int rc, const char *op) {
  TORCH_CHECK(rc == PYRONN_OK, op, " failed (status ", rc, "): ", pyronn_last_error());
}

pyronn_stream_t cur_stream(const at::Tensor &t) {
  return reinterpret_cast<pyronn_stream_t>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

void check_data(const at::Tensor &t, int64_t dim, const char *op, const char *name) {
  TORCH_CHECK(t.is_cuda(), op, ": ", name, " must be a GPU tensor (no CPU kernel is registered)");
  TORCH_CHECK(t.dim() == dim, op, ": ", name, " must have ", dim, " dims [batch, ...], got ", t.dim());
}

// ------------------------------------------------------------------------ cone

at::Tensor cone_projection3d(const at::Tensor &volume, at::IntArrayRef projection_shape,
                             const at::Tensor &volume_origin, const at::Tensor &volume_spacing,
                             const at::Tensor &projection_matrices, const at::Tensor &step_size,
                             const at::Tensor &projection_multiplier, bool hardware_interp) {
  const char *op = "cone_projection3d";
  check_data(volume, 4, op, "volume");
  TORCH_CHECK(projection_shape.size() == 3, op, ": projection_shape must be (n, det_h, det_w)");
  c10::DeviceGuard guard(volume.device());
  const at::Tensor vol = volume.contiguous();
  const int64_t B = vol.size(0), D = vol.size(1), H = vol.size(2), W = vol.size(3);
  const int64_t n = projection_shape[0], Hd = projection_shape[1], Wd = projection_shape[2];
  const auto dev = vol.device();
  const at::Tensor org = geom(volume_origin, B, {3}, op, "volume_origin", dev);
  const at::Tensor sp = geom(volume_spacing, B, {3}, op, "volume_spacing", dev);
  const at::Tensor P = geom(projection_matrices, B, {n, 3, 4}, op, "projection_matrices", dev);
  const at::Tensor st = geom(step_size, B, {1}, op, "step_size", dev);
  const at::Tensor mu = geom(projection_multiplier, B, {1}, op, "projection_multiplier", dev);
  const int dt = dtype_code(vol, op);
  at::Tensor out = at::empty({B, n, Hd, Wd}, vol.options().dtype(at::kFloat));
  if (empty_output(out)) return out;
  const size_t ws = pyronn_cone_projection3d_workspace_size((int)B, (int)D, (int)H, (int)W, (int)n, dt);
  TORCH_CHECK(ws > 0, op, ": invalid shape");
  at::Tensor work = at::empty({(int64_t)ws}, vol.options().dtype(at::kByte));
  check_rc(pyronn_cone_projection3d(vol.data_ptr(), dt, (int)B, (int)D, (int)H, (int)W,
                                    org.data_ptr<float>(), sp.data_ptr<float>(), P.data_ptr<float>(),
                                    st.data_ptr<float>(), mu.data_ptr<float>(), (int)n, (int)Hd,
                                    (int)Wd, hardware_interp ? 1 : 0, out.data_ptr<float>(),
                                    work.data_ptr(), ws, cur_stream(vol)),
           op);
  return out;
}

at::Tensor cone_backprojection3d(const at::Tensor &sinogram, at::IntArrayRef volume_shape,
                                 const at::Tensor &volume_origin, const at::Tensor &volume_spacing,
                                 const at::Tensor &projection_matrices, const at::Tensor &step_size,
                                 const at::Tensor &projection_multiplier, bool hardware_interp) {
  const char *op = "cone_backprojection3d";
  check_data(sinogram, 4, op, "sinogram");
  TORCH_CHECK(volume_shape.size() == 3, op, ": volume_shape must be (D, H, W)");
  c10::DeviceGuard guard(sinogram.device());
  const at::Tensor sino = sinogram.contiguous();
  const int64_t B = sino.size(0), n = sino.size(1), Hd = sino.size(2), Wd = sino.size(3);
  const int64_t D = volume_shape[0], H = volume_shape[1], W = volume_shape[2];
  const auto dev = sino.device();
  const at::Tensor org = geom(volume_origin, B, {3}, op, "volume_origin", dev);
  const at::Tensor sp = geom(volume_spacing, B, {3}, op, "volume_spacing", dev);
  const at::Tensor P = geom(projection_matrices, B, {n, 3, 4}, op, "projection_matrices", dev);
  const at::Tensor st = geom(step_size, B, {1}, op, "step_size", dev);
  const at::Tensor mu = geom(projection_multiplier, B, {1}, op, "projection_multiplier", dev);
  const int dt = dtype_code(sino, op);
  at::Tensor out = at::empty({B, D, H, W}, sino.options().dtype(at::kFloat));
  if (empty_output(out)) return out;
  const size_t ws = pyronn_cone_backprojection3d_workspace_size((int)B, (int)n, (int)Hd, (int)Wd, dt);
  TORCH_CHECK(ws > 0, op, ": invalid shape");
  at::Tensor work = at::empty({(int64_t)ws}, sino.options().dtype(at::kByte));
  check_rc(pyronn_cone_backprojection3d(sino.data_ptr(), dt, (int)B, (int)n, (int)Hd, (int)Wd,
                                        org.data_ptr<float>(), sp.data_ptr<float>(),
                                        P.data_ptr<float>(), st.data_ptr<float>(),
                                        mu.data_ptr<float>(), (int)D, (int)H, (int)W,
                                        hardware_interp ? 1 : 0, out.data_ptr<float>(),
                                        work.data_ptr(), ws, cur_stream(sino)),
           op);
  return out;
}

// ------------------------------------------------------------------------ 2D

struct Geo2D {
  at::Tensor vo, dor, vs, ds, sid, sdd, rays;
};

Geo2D geo2d(int64_t B, int64_t n, const at::Tensor &vo, const at::Tensor &dor, const at::Tensor &vs,
            const at::Tensor &ds, const at::Tensor *sid, const at::Tensor *sdd,
            const at::Tensor &rays, const char *op, const at::Device &dev) {
  Geo2D g;
  g.vo = geom(vo, B, {2}, op, "volume_origin", dev);
  g.dor = geom(dor, B, {1}, op, "detector_origin", dev);
  g.vs = geom(vs, B, {2}, op, "volume_spacing", dev);
  g.ds = geom(ds, B, {1}, op, "detector_spacing", dev);
  if (sid) g.sid = geom(*sid, B, {1}, op, "source_2_isocenter_distance", dev);
  if (sdd) g.sdd = geom(*sdd, B, {1}, op, "source_2_detector_distance", dev);
  g.rays = geom(rays, B, {n, 2}, op, "ray_vectors", dev);
  return g;
}

at::Tensor proj2d(bool fan, const at::Tensor &volume, at::IntArrayRef projection_shape,
                  const at::Tensor &vo, const at::Tensor &dor, const at::Tensor &vs,
                  const at::Tensor &ds, const at::Tensor *sid, const at::Tensor *sdd,
                  const at::Tensor &rays, double step_size) {
  const char *op = fan ? "fan_projection2d" : "parallel_projection2d";
  check_data(volume, 3, op, "volume");
  TORCH_CHECK(projection_shape.size() == 2, op, ": projection_shape must be (n, det)");
  c10::DeviceGuard guard(volume.device());
  const at::Tensor vol = volume.contiguous();
  const int64_t B = vol.size(0), H = vol.size(1), W = vol.size(2);
  const int64_t n = projection_shape[0], det = projection_shape[1];
  Geo2D g = geo2d(B, n, vo, dor, vs, ds, sid, sdd, rays, op, vol.device());
  const int dt = dtype_code(vol, op);
  at::Tensor out = at::empty({B, n, det}, vol.options().dtype(at::kFloat));
  if (empty_output(out)) return out;
  int rc;
  if (fan)
    rc = pyronn_fan_projection2d(vol.data_ptr(), dt, (int)B, (int)H, (int)W, g.vo.data_ptr<float>(),
                                 g.dor.data_ptr<float>(), g.vs.data_ptr<float>(), g.ds.data_ptr<float>(),
                                 g.sid.data_ptr<float>(), g.sdd.data_ptr<float>(),
                                 g.rays.data_ptr<float>(), (int)n, (int)det, (float)step_size,
                                 out.data_ptr<float>(), cur_stream(vol));
  else
    rc = pyronn_parallel_projection2d(vol.data_ptr(), dt, (int)B, (int)H, (int)W,
                                      g.vo.data_ptr<float>(), g.dor.data_ptr<float>(),
                                      g.vs.data_ptr<float>(), g.ds.data_ptr<float>(),
                                      g.rays.data_ptr<float>(), (int)n, (int)det, (float)step_size,
                                      out.data_ptr<float>(), cur_stream(vol));
  check_rc(rc, op);
  return out;
}

at::Tensor back2d(bool fan, const at::Tensor &sinogram, at::IntArrayRef volume_shape,
                  const at::Tensor &vo, const at::Tensor &dor, const at::Tensor &vs,
                  const at::Tensor &ds, const at::Tensor *sid, const at::Tensor *sdd,
                  const at::Tensor &rays) {
  const char *op = fan ? "fan_backprojection2d" : "parallel_backprojection2d";
  check_data(sinogram, 3, op, "sinogram");
  TORCH_CHECK(volume_shape.size() == 2, op, ": volume_shape must be (H, W)");
  c10::DeviceGuard guard(sinogram.device());
  const at::Tensor sino = sinogram.contiguous();
  const int64_t B = sino.size(0), n = sino.size(1), det = sino.size(2);
  const int64_t H = volume_shape[0], W = volume_shape[1];
  Geo2D g = geo2d(B, n, vo, dor, vs, ds, sid, sdd, rays, op, sino.device());
  const int dt = dtype_code(sino, op);
  at::Tensor out = at::empty({B, H, W}, sino.options().dtype(at::kFloat));
  if (empty_output(out)) return out;
  int rc;
  if (fan)
    rc = pyronn_fan_backprojection2d(sino.data_ptr(), dt, (int)B, (int)n, (int)det,
                                     g.vo.data_ptr<float>(), g.dor.data_ptr<float>(),
                                     g.vs.data_ptr<float>(), g.ds.data_ptr<float>(),
                                     g.sid.data_ptr<float>(), g.sdd.data_ptr<float>(),
                                     g.rays.data_ptr<float>(), (int)H, (int)W,
                                     out.data_ptr<float>(), cur_stream(sino));
  else
    rc = pyronn_parallel_backprojection2d(sino.data_ptr(), dt, (int)B, (int)n, (int)det,
                                          g.vo.data_ptr<float>(), g.dor.data_ptr<float>(),
                                          g.vs.data_ptr<float>(), g.ds.data_ptr<float>(),
                                          g.rays.data_ptr<float>(), (int)H, (int)W,
                                          out.data_ptr<float>(), cur_stream(sino));
  check_rc(rc, op);
  return out;
}

at::Tensor parallel_projection2d(const at::Tensor &volume, at::IntArrayRef projection_shape,
                                 const at::Tensor &volume_origin, const at::Tensor &detector_origin,
                                 const at::Tensor &volume_spacing, const at::Tensor &detector_spacing,
                                 const at::Tensor &ray_vectors, double step_size) {
  return proj2d(false, volume, projection_shape, volume_origin, detector_origin, volume_spacing,
                detector_spacing, nullptr, nullptr, ray_vectors, step_size);
}

at::Tensor parallel_backprojection2d(const at::Tensor &sinogram, at::IntArrayRef volume_shape,
                                     const at::Tensor &volume_origin,
                                     const at::Tensor &detector_origin,
                                     const at::Tensor &volume_spacing,
                                     const at::Tensor &detector_spacing,
                                     const at::Tensor &ray_vectors) {
  return back2d(false, sinogram, volume_shape, volume_origin, detector_origin, volume_spacing,
                detector_spacing, nullptr, nullptr, ray_vectors);
}

at::Tensor fan_projection2d(const at::Tensor &volume, at::IntArrayRef projection_shape,
                            const at::Tensor &volume_origin, const at::Tensor &detector_origin,
                            const at::Tensor &volume_spacing, const at::Tensor &detector_spacing,
                            const at::Tensor &source_2_isocenter_distance,
                            const at::Tensor &source_2_detector_distance,
                            const at::Tensor &central_ray_vectors, double step_size) {
  return proj2d(true, volume, projection_shape, volume_origin, detector_origin, volume_spacing,
                detector_spacing, &source_2_isocenter_distance, &source_2_detector_distance,
                central_ray_vectors, step_size);
}

at::Tensor fan_backprojection2d(const at::Tensor &sinogram, at::IntArrayRef volume_shape,
                                const at::Tensor &volume_origin, const at::Tensor &detector_origin,
                                const at::Tensor &volume_spacing, const at::Tensor &detector_spacing,
                                const at::Tensor &source_2_isocenter_distance,
                                const at::Tensor &source_2_detector_distance,
                                const at::Tensor &central_ray_vectors) {
  return back2d(true, sinogram, volume_shape, volume_origin, detector_origin, volume_spacing,
                detector_spacing, &source_2_isocenter_distance, &source_2_detector_distance,
                central_ray_vectors);
}

}  // namespace

TORCH_LIBRARY(pyronn_layers, m) {
  m.def("parallel_projection2d(Tensor volume, int[] projection_shape, Tensor volume_origin, "
        "Tensor detector_origin, Tensor volume_spacing, Tensor detector_spacing, "
        "Tensor ray_vectors, float step_size=0.2) -> Tensor");
  m.def("parallel_backprojection2d(Tensor sinogram, int[] volume_shape, Tensor volume_origin, "
        "Tensor detector_origin, Tensor volume_spacing, Tensor detector_spacing, "
        "Tensor ray_vectors) -> Tensor");
  m.def("fan_projection2d(Tensor volume, int[] projection_shape, Tensor volume_origin, "
        "Tensor detector_origin, Tensor volume_spacing, Tensor detector_spacing, "
        "Tensor source_2_isocenter_distance, Tensor source_2_detector_distance, "
        "Tensor central_ray_vectors, float step_size=0.2) -> Tensor");
  m.def("fan_backprojection2d(Tensor sinogram, int[] volume_shape, Tensor volume_origin, "
        "Tensor detector_origin, Tensor volume_spacing, Tensor detector_spacing, "
        "Tensor source_2_isocenter_distance, Tensor source_2_detector_distance, "
        "Tensor central_ray_vectors) -> Tensor");
  m.def("cone_projection3d(Tensor volume, int[] projection_shape, Tensor volume_origin, "
        "Tensor volume_spacing, Tensor projection_matrices, Tensor step_size, "
        "Tensor projection_multiplier, bool hardware_interp=False) -> Tensor");
  m.def("cone_backprojection3d(Tensor sinogram, int[] volume_shape, Tensor volume_origin, "
        "Tensor volume_spacing, Tensor projection_matrices, Tensor step_size, "
        "Tensor projection_multiplier, bool hardware_interp=False) -> Tensor");
}

TORCH_LIBRARY_IMPL(pyronn_layers, CUDA, m) {
  m.impl("parallel_projection2d", &parallel_projection2d);
  m.impl("parallel_backprojection2d", &parallel_backprojection2d);
  m.impl("fan_projection2d", &fan_projection2d);
  m.impl("fan_backprojection2d", &fan_backprojection2d);
  m.impl("cone_projection3d", &cone_projection3d);
  m.impl("cone_backprojection3d", &cone_backprojection3d);
}
