// Discrete-adjoint VJP entry points (placeholder until the adjoint kernels land).
#include <hip/hip_runtime.h>

#include "../../include/gravlume.h"

extern "C" {

gl_status gl_vjp_forward(gl_trace_plan*, uint32_t, double, uint32_t, const gl_ray_outputs*, void*,
                         gl_vjp_state** out_state) {
  if (out_state) *out_state = nullptr;
  return GL_ERR_INVALID_ARGUMENT;
}

gl_status gl_vjp_backward(gl_vjp_state*, const float*, double*, float*, void*) {
  return GL_ERR_INVALID_ARGUMENT;
}

gl_status gl_vjp_destroy(gl_vjp_state*) { return GL_OK; }

}  // extern "C"
