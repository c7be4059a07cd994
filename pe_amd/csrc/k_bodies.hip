// Per-body kernels: AABBs, the velocity predictor (pe's "body caching") and the
// semi-implicit position integrator.  One thread per body; all HBM-bound.
#include "step_ops.hpp"

namespace pe_amd {

// AABB of every body from its current state.
__global__ void k_aabb(Bodies B) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B.n) update_aabb(B, i);
}

// Body caching (predicted velocities, solver body records); see predict_body.
__global__ void k_predict(Bodies B, d3 gravity, double dt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B.n) predict_body(B, i, gravity, dt);
}

// Velocity corrections applied, positions integrated; see integrate_body.
template <class Trig>
__global__ void k_integrate(Bodies B, double dt, double damping) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B.n) integrate_body<Trig>(B, i, dt, damping);
}

// Kinetic energy 0.5*m*v^2 + 0.5*w.(I_w w) of the non-fixed bodies, one partial per block.
__global__ void k_kinetic(Bodies B, double* partial) {
  __shared__ double red[256];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0;
  if (i < B.n && !B.fixed[i]) {
    const m33 R = ld9(B.R, i);
    const m33 Iwf = mat_mul(mat_mul(R, ld9(B.Ib, i)), mat_trans(R));
    const d3 v = ld3(B.v, i), w = ld3(B.w, i);
    const double m = 1.0 / B.inv_mass[i];
    e = 0.5 * m * dot3(v, v) + 0.5 * dot3(w, mat_vec(Iwf, w));
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

void launch_aabb(const Bodies& B, hipStream_t s) {
  if (B.n) hipLaunchKernelGGL(k_aabb, dim3((B.n + 255) / 256), dim3(256), 0, s, B);
}
void launch_predict(const Bodies& B, d3 g, double dt, hipStream_t s) {
  if (B.n) hipLaunchKernelGGL(k_predict, dim3((B.n + 255) / 256), dim3(256), 0, s, B, g, dt);
}
void launch_integrate(const Bodies& B, double dt, double damping, bool exact_trig, hipStream_t s) {
  if (!B.n) return;
  if (exact_trig)
    hipLaunchKernelGGL(k_integrate<TrigExact>, dim3((B.n + 255) / 256), dim3(256), 0, s, B, dt, damping);
  else
    hipLaunchKernelGGL(k_integrate<TrigFast>, dim3((B.n + 255) / 256), dim3(256), 0, s, B, dt, damping);
}

// Device evaluation of one trig function over an array (checks of the device trig against
// the host libm): fn 0 sin, 1 cos, 2 atan2(y, x), 3 acos, 4 / 5 sin / cos from sincos.
template <class Trig>
__global__ void k_eval_trig(int fn, size_t n, const double* x, const double* y, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r;
  switch (fn) {
    case 0: r = Trig::sin(x[i]); break;
    case 1: r = Trig::cos(x[i]); break;
    case 2: r = Trig::atan2(y[i], x[i]); break;
    case 3: r = Trig::acos(x[i]); break;
    default: {
      double sn, cs;
      Trig::sincos(x[i], sn, cs);
      r = fn == 4 ? sn : cs;
    }
  }
  out[i] = r;
}

void launch_eval_trig(int fn, bool exact, size_t n, const double* x, const double* y, double* out, hipStream_t s) {
  if (!n) return;
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
  if (exact) hipLaunchKernelGGL(k_eval_trig<TrigExact>, g, b, 0, s, fn, n, x, y, out);
  else hipLaunchKernelGGL(k_eval_trig<TrigFast>, g, b, 0, s, fn, n, x, y, out);
}
uint32_t launch_kinetic(const Bodies& B, double* partial, hipStream_t s) {
  const uint32_t nb = (B.n + 255) / 256;
  if (B.n) hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(256), 0, s, B, partial);
  return nb;
}

}  // namespace pe_amd
