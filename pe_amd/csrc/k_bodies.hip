// Per-body kernels: AABBs, the velocity predictor (pe's "body caching") and the
// semi-implicit position integrator.  One thread per body; all HBM-bound.
#include "kernels.hpp"

namespace pe_amd {

// AABB of every body from its current state.
__global__ void k_aabb(Bodies B) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n) return;
  const m33 R = ld9(B.R, i);
  double o[6];
  body_aabb(B.shape[i], ld3(B.pos, i), R, B.param + 3 * (size_t)i, ld3(B.pnormal, i), B.pd[i], o);
#pragma unroll
  for (int k = 0; k < 6; ++k) B.aabb[6 * (size_t)i + k] = o[k];
}

// Body caching, HCSITS.h:1834-1878 with initializeVelocityCorrections, HCSITS.h:4094-4109.
// Writes the solver body record (predicted v, w; inverse mass; R*Iinv*R^T) and the
// initial velocity corrections from external forces; resets the forces.
__global__ void k_predict(Bodies B, d3 gravity, double dt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n) return;
  const m33 R = ld9(B.R, i);
  const m33 Rt = mat_trans(R);
  const m33 Iw = mat_mul(mat_mul(R, ld9(B.Ibinv, i)), Rt);  // RigidBody::getInvInertia
  const double im = B.inv_mass[i];
  d3 v = ld3(B.v, i), w = ld3(B.w, i);
  double* q = B.dvw + (size_t)kDVW * i;
  if (!B.fixed[i]) {
    const d3 F = ld3(B.force, i), T = ld3(B.torque, i);
    const d3 dv = (im * dt) * F;
    const d3 dw = dt * mat_vec(Iw, T);
    q[0] = dv.x; q[1] = dv.y; q[2] = dv.z; q[3] = dw.x; q[4] = dw.y; q[5] = dw.z;
    const m33 Iwf = mat_mul(mat_mul(R, ld9(B.Ib, i)), Rt);  // RigidBody::getInertia
    const d3 v0 = v, w0 = w;
    v = v0 + gravity * dt;
    w = w0 + dt * mat_vec(Iw, cross3(mat_vec(Iwf, w0), w0));
  }
  st3(B.force, i, mk3(0, 0, 0));
  st3(B.torque, i, mk3(0, 0, 0));
  double* s = B.sb + (size_t)kSB * i;
  s[SB_V] = v.x; s[SB_V + 1] = v.y; s[SB_V + 2] = v.z;
  s[SB_W] = w.x; s[SB_W + 1] = w.y; s[SB_W + 2] = w.z;
  s[SB_IM] = im;
#pragma unroll
  for (int k = 0; k < 9; ++k) s[SB_IW + k] = Iw.a[k];
}

// Apply the velocity corrections and integrate positions: resolveContacts' final
// loop (HCSITS.h:1954-2004) and integratePositions (HCSITS.h:4135-4184).
template <class Trig>
__global__ void k_integrate(Bodies B, double dt, double damping) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n) return;
  const double* s = B.sb + (size_t)kSB * i;
  const double* q = B.dvw + (size_t)kDVW * i;
  d3 v = mk3(s[SB_V], s[SB_V + 1], s[SB_V + 2]) + mk3(q[0], q[1], q[2]);
  d3 w = mk3(s[SB_W], s[SB_W + 1], s[SB_W + 2]) + mk3(q[3], q[4], q[5]);
  d3 gp = ld3(B.pos, i) + v * dt;
  const d3 phi = w * dt;
  const qd dq = quat_axis_angle<Trig>(phi, len3(phi));
  const qd qn = quat_mul(dq, ld4(B.q, i));
  const m33 R = quat_to_R(qn);
  if (damping < 1.0) {
    const double drag = pow(damping, dt);
    v = v * drag;
    w = w * drag;
  }
  st3(B.pos, i, gp);
#pragma unroll
  for (int k = 0; k < 4; ++k) B.q[4 * (size_t)i + k] = qn.v[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) B.R[9 * (size_t)i + k] = R.a[k];
  st3(B.v, i, v);
  st3(B.w, i, w);
  d3 nrm = ld3(B.pnormal, i);
  double pd = B.pd[i];
  const int shape = B.shape[i];
  if (shape == kShapePlane) {
    nrm = quat_rotate(qn, mk3(0, 0, 1));
    pd = dot3(nrm, gp);
    st3(B.pnormal, i, nrm);
    B.pd[i] = pd;
  }
  double o[6];
  body_aabb(shape, gp, R, B.param + 3 * (size_t)i, nrm, pd, o);
#pragma unroll
  for (int k = 0; k < 6; ++k) B.aabb[6 * (size_t)i + k] = o[k];
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
// the host libm): fn 0 sin, 1 cos, 2 atan2(y, x), 3 acos.
template <class Trig>
__global__ void k_eval_trig(int fn, size_t n, const double* x, const double* y, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r;
  switch (fn) {
    case 0: r = Trig::sin(x[i]); break;
    case 1: r = Trig::cos(x[i]); break;
    case 2: r = Trig::atan2(y[i], x[i]); break;
    default: r = Trig::acos(x[i]); break;
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
