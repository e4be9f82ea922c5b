// Thread-local error reporting of the C ABI (no exit(), no exceptions across
// the boundary; the reference aborted the process on any CUDA error,
// cc/kernels/cone_projector_3D_CudaKernel.cu.cc:32-40).
#include "common.h"

namespace pyronn {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

int hip_fail(hipError_t e, const char *what) {
  set_error("%s: HIP error %d (%s)", what, (int)e, hipGetErrorString(e));
  return PYRONN_ERR_HIP;
}

}  // namespace pyronn

extern "C" const char *pyronn_last_error(void) { return pyronn::g_last_error.c_str(); }

extern "C" int pyronn_version(void) { return 100; }
