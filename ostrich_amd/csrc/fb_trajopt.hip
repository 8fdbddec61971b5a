// fb_trajopt.hip -- Adam on a (steps, worlds, dofs) control trajectory, restating
// optim/trajectory_adam.py:6-53 (adam_step_kernel_3d): per element, the raw gradient is
// clamped to [-clip, clip] (clip > 0), the biased moments are updated, bias-corrected
// with beta^(t+1), and the parameter moves by -lr mhat / (sqrt(vhat) + eps).  Only the
// dof slice [dof_offset, dof_offset + num_dofs) of each (step, world) is touched.

#include <hip/hip_runtime.h>

#include "fb_solver.h"

namespace {

__global__ void __launch_bounds__(256) trajectory_adam_kernel(float* __restrict__ params,
                                                              const float* __restrict__ grad,
                                                              float* __restrict__ m, float* __restrict__ v,
                                                              int rows, int dofs, int dof_offset, int num_dofs,
                                                              float lr, float beta1, float beta2, float eps,
                                                              float clip, float bc1, float bc2) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)rows * num_dofs) return;
  const long long i = (e / num_dofs) * dofs + dof_offset + e % num_dofs;
  float g = grad[i];
  if (clip > 0.f) g = fminf(fmaxf(g, -clip), clip);
  const float mn = beta1 * m[i] + (1.f - beta1) * g;
  const float vn = beta2 * v[i] + (1.f - beta2) * g * g;
  m[i] = mn;
  v[i] = vn;
  const float mhat = mn / bc1, vhat = vn / bc2;
  params[i] = params[i] - lr * mhat / (sqrtf(vhat) + eps);
}

}  // namespace

extern "C" int ofb_trajectory_adam_step(float* params, const float* grad, float* m, float* v, int32_t steps,
                                        int32_t worlds, int32_t dofs, int32_t dof_offset, int32_t num_dofs,
                                        float lr, float beta1, float beta2, float eps, float clip_grad, int32_t t,
                                        void* stream) {
  if (!params || !grad || !m || !v) {
    ofb::set_error("trajectory Adam: null array");
    return -1;
  }
  if (steps < 0 || worlds < 0 || dofs < 1 || dof_offset < 0 || num_dofs < 0 || dof_offset + num_dofs > dofs ||
      t < 0) {
    ofb::set_error("trajectory Adam: dof slice [%d, %d) outside %d dofs (or negative sizes)", dof_offset,
                   dof_offset + num_dofs, dofs);
    return -2;
  }
  const long long n = (long long)steps * worlds * num_dofs;
  if (n == 0) return 0;
  // bias corrections 1 - beta^(t+1) (trajectory_adam.py:48-49)
  const float bc1 = 1.f - powf(beta1, (float)t + 1.f), bc2 = 1.f - powf(beta2, (float)t + 1.f);
  const int T = 256;
  hipLaunchKernelGGL(trajectory_adam_kernel, dim3((unsigned)((n + T - 1) / T)), dim3(T), 0, (hipStream_t)stream,
                     params, grad, m, v, steps * worlds, dofs, dof_offset, num_dofs, lr, beta1, beta2, eps, clip_grad,
                     bc1, bc2);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    ofb::set_error("trajectory Adam launch failed: %s", hipGetErrorString(err));
    return -4;
  }
  return 0;
}
