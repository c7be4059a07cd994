// Kernels of the adjoint time step (spheres and planes, relaxation models 0, 1 and 3).
//
// For one recorded step (contacts in sweep order with their converged impulses, positions at
// the start of the step, solver body records, post-solve velocities):
//   k_adj_jacobian  one thread per (contact, input direction): contact_map() in dual numbers
//                   gives one column of the contact's 15 x 22 Jacobian;
//   k_adj_prepare   per contact, the 3 x 3 matrix of the transposed local system and its inverse;
//   k_adj_init      body adjoints alpha = adjoint of the post-solve velocities;
//   k_adj_sweep     one colour of a Gauss-Seidel sweep on the transposed (adjoint) system;
//   k_adj_assemble  position and friction adjoints from the converged (alpha, lambda).
#include "adjoint.hpp"
#include "kernels.hpp"

namespace pe_amd {

using ad::ContactConst;
using ad::Dual;
using ad::kAdjIn;
using ad::kAdjOut;

__device__ __forceinline__ ContactConst make_const(const AdjTape& T, const AdjConsts& A, uint32_t s, uint32_t& i1,
                                                   uint32_t& i2, bool& f1, bool& f2) {
  const uint32_t e1 = T.b1[s], e2 = T.b2[s];
  i1 = e1 & ~kFixedBit;
  i2 = e2 & ~kFixedBit;
  f1 = (e1 & kFixedBit) != 0;
  f2 = (e2 & kFixedBit) != 0;
  ContactConst K;
  K.rad1 = A.param[3 * (size_t)i1];
  K.rad2 = A.param[3 * (size_t)i2];
  const double* s1 = T.sb + (size_t)kSB * i1;
  const double* s2 = T.sb + (size_t)kSB * i2;
  K.im1 = s1[SB_IM];
  K.im2 = s2[SB_IM];
#pragma unroll
  for (int k = 0; k < 9; ++k) { K.I1[k] = s1[SB_IW + k]; K.I2[k] = s2[SB_IW + k]; }
  K.plane = A.shape[i2] == kShapePlane ? 1 : 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    K.pn[k] = A.pnormal[3 * (size_t)i2 + k];
    K.xp[k] = A.ppos[3 * (size_t)i2 + k];
  }
  K.pd = A.pd[i2];
  K.erp = A.erp;
  K.dtinv = A.dtinv;
  return K;
}

template <int MODEL>
__global__ void k_adj_jacobian(AdjTape T, AdjConsts A, double* jac) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = (uint32_t)(g / kAdjIn), j = (uint32_t)(g % kAdjIn);
  if (s >= T.nc) return;
  uint32_t i1, i2;
  bool f1, f2;
  const ContactConst K = make_const(T, A, s, i1, i2, f1, f2);
  Dual z[kAdjIn];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    z[k] = ad::mkd(T.pos[3 * (size_t)i1 + k]);
    z[3 + k] = ad::mkd(T.pos[3 * (size_t)i2 + k]);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    z[6 + k] = ad::mkd(f1 ? 0.0 : T.vel[6 * (size_t)i1 + k]);
    z[12 + k] = ad::mkd(f2 ? 0.0 : T.vel[6 * (size_t)i2 + k]);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) z[18 + k] = ad::mkd(T.p[3 * (size_t)s + k]);
  z[21] = ad::mkd(T.mu[s]);
  z[j].d = 1.0;
  Dual out[kAdjOut];
  ad::contact_map<Dual, MODEL>(K, z, out);
  double* J = jac + ((size_t)s * kAdjIn + j) * kAdjOut;
#pragma unroll
  for (int k = 0; k < kAdjOut; ++k) J[k] = out[k].d;
}

// Jacobian entry d out[k] / d z[j] of contact s.
__device__ __forceinline__ double jget(const double* jac, uint32_t s, int j, int k) {
  return jac[((size_t)s * kAdjIn + j) * kAdjOut + k];
}

// Mc = P_p^T - Q_p^T P_U^T (P_p = dE2/dp, Q_p = d(-contrib)/dp, P_U = dE2/dU), inverted.
__global__ void k_adj_prepare(uint32_t nc, const double* jac, double* minv) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  m33 M;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double v = jget(jac, s, 18 + a, b);                    // P_p[b][a]
      for (int m = 0; m < 12; ++m) v -= jget(jac, s, 18 + a, 3 + m) * jget(jac, s, 6 + m, b);
      M.a[3 * a + b] = v;
    }
  const m33 Mi = mat_inv(M);
#pragma unroll
  for (int k = 0; k < 9; ++k) minv[9 * (size_t)s + k] = Mi.a[k];
}

// alpha_b = (vbar + dt * xbar, wbar) for non-fixed bodies, 0 for fixed ones.
__global__ void k_adj_init(uint32_t n, const uint8_t* fixed, const double* xbar, const double* vbar,
                           const double* wbar, double dt, double* alpha) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  double* a = alpha + 6 * (size_t)b;
  if (fixed[b]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = 0.0;
    return;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a[k] = vbar[3 * (size_t)b + k] + dt * xbar[3 * (size_t)b + k];
    a[3 + k] = wbar[3 * (size_t)b + k];
  }
}

__device__ __forceinline__ void wave_max_u64(double v, unsigned long long* out) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(u, d, 64);
    u = o > u ? o : u;
  }
  if ((threadIdx.x & 63) == 0 && u) atomicMax(out, u);
}

// One colour of the adjoint sweep: contacts [begin, end) share no non-fixed body.
__global__ void k_adj_sweep(AdjTape T, uint32_t begin, uint32_t end, const double* jac, const double* minv,
                            double* alpha, double* lambda, unsigned long long* delta) {
  const uint32_t s = begin + blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (s < end) {
    const uint32_t e1 = T.b1[s], e2 = T.b2[s];
    const uint32_t i1 = e1 & ~kFixedBit, i2 = e2 & ~kFixedBit;
    const bool f1 = (e1 & kFixedBit) != 0, f2 = (e2 & kFixedBit) != 0;
    double a[12];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      a[k] = f1 ? 0.0 : alpha[6 * (size_t)i1 + k];
      a[6 + k] = f2 ? 0.0 : alpha[6 * (size_t)i2 + k];
    }
    double lo[3], ln[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) lo[k] = lambda[3 * (size_t)s + k];
    // remove the contact's own term: a += P_U^T lambda_old
#pragma unroll
    for (int m = 0; m < 12; ++m)
      a[m] += jget(jac, s, 6 + m, 0) * lo[0] + jget(jac, s, 6 + m, 1) * lo[1] + jget(jac, s, 6 + m, 2) * lo[2];
    double rhs[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int m = 0; m < 12; ++m) v -= jget(jac, s, 18 + c, 3 + m) * a[m];
      rhs[c] = v;
    }
    const double* Mi = minv + 9 * (size_t)s;
#pragma unroll
    for (int c = 0; c < 3; ++c) ln[c] = Mi[3 * c] * rhs[0] + Mi[3 * c + 1] * rhs[1] + Mi[3 * c + 2] * rhs[2];
#pragma unroll
    for (int m = 0; m < 12; ++m)
      a[m] -= jget(jac, s, 6 + m, 0) * ln[0] + jget(jac, s, 6 + m, 1) * ln[1] + jget(jac, s, 6 + m, 2) * ln[2];
    if (!f1)
#pragma unroll
      for (int k = 0; k < 6; ++k) alpha[6 * (size_t)i1 + k] = a[k];
    if (!f2)
#pragma unroll
      for (int k = 0; k < 6; ++k) alpha[6 * (size_t)i2 + k] = a[6 + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lambda[3 * (size_t)s + k] = ln[k];
      d = fmax(d, fabs(ln[k] - lo[k]));
    }
  }
  wave_max_u64(d, delta);
}

// xbar += -(Q_x^T alpha_c + P_x^T lambda_c) per contact; mubar += -P_mu . lambda.
__global__ void k_adj_assemble(AdjTape T, const double* jac, const double* alpha, const double* lambda, double* xbar,
                               double* mubar) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  double mg = 0.0;
  if (s < T.nc) {
    const uint32_t e1 = T.b1[s], e2 = T.b2[s];
    const uint32_t i1 = e1 & ~kFixedBit, i2 = e2 & ~kFixedBit;
    const bool f1 = (e1 & kFixedBit) != 0, f2 = (e2 & kFixedBit) != 0;
    double a[12];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      a[k] = f1 ? 0.0 : alpha[6 * (size_t)i1 + k];
      a[6 + k] = f2 ? 0.0 : alpha[6 * (size_t)i2 + k];
    }
    const double l0 = lambda[3 * (size_t)s], l1 = lambda[3 * (size_t)s + 1], l2 = lambda[3 * (size_t)s + 2];
    double xg[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double v = jget(jac, s, j, 0) * l0 + jget(jac, s, j, 1) * l1 + jget(jac, s, j, 2) * l2;
      for (int m = 0; m < 12; ++m) v += jget(jac, s, j, 3 + m) * a[m];
      xg[j] = -v;
    }
    if (!f1)
      for (int k = 0; k < 3; ++k) atomicAdd(&xbar[3 * (size_t)i1 + k], xg[k]);
    if (!f2)
      for (int k = 0; k < 3; ++k) atomicAdd(&xbar[3 * (size_t)i2 + k], xg[3 + k]);
    mg = -(jget(jac, s, 21, 0) * l0 + jget(jac, s, 21, 1) * l1 + jget(jac, s, 21, 2) * l2);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) mg += __shfl_xor(mg, d, 64);
  if ((threadIdx.x & 63) == 0 && mg != 0.0) atomicAdd(mubar, mg);
}

// vbar = alpha_v, wbar = alpha_w (the predictor passes velocities through for spheres).
__global__ void k_adj_finish(uint32_t n, const double* alpha, double* vbar, double* wbar) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    vbar[3 * (size_t)b + k] = alpha[6 * (size_t)b + k];
    wbar[3 * (size_t)b + k] = alpha[6 * (size_t)b + 3 + k];
  }
}

// Rotation matrices from quaternions and AABBs (restoring a checkpoint on the device).
__global__ void k_adj_restore_R(uint32_t n, const double* q, double* R) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  qd qq;
#pragma unroll
  for (int k = 0; k < 4; ++k) qq.v[k] = q[4 * (size_t)i + k];
  const m33 m = quat_to_R(qq);
#pragma unroll
  for (int k = 0; k < 9; ++k) R[9 * (size_t)i + k] = m.a[k];
}

// Body-major copy of (v, w) into 6 doubles per body.
__global__ void k_adj_pack_vel(uint32_t n, const double* v, const double* w, double* vel) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    vel[6 * (size_t)i + k] = v[3 * (size_t)i + k];
    vel[6 * (size_t)i + 3 + k] = w[3 * (size_t)i + k];
  }
}

// Impulses of solver slots as 3 doubles per contact (from the field-major solver arrays).
__global__ void k_adj_pack_p(uint32_t nc, const double* f, size_t cap, double* p, double* mu) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) p[3 * (size_t)s + k] = f[(CF_P + k) * cap + s];
  mu[s] = f[CF_MU * cap + s];
}

// ---------------------------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------------------------
static inline dim3 blocks(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

void adj_jacobian(int model, const AdjTape& T, const AdjConsts& A, double* jac, hipStream_t s) {
  if (!T.nc) return;
  const size_t n = (size_t)T.nc * kAdjIn;
  if (model == 0) hipLaunchKernelGGL(k_adj_jacobian<0>, blocks(n), dim3(256), 0, s, T, A, jac);
  else if (model == 1) hipLaunchKernelGGL(k_adj_jacobian<1>, blocks(n), dim3(256), 0, s, T, A, jac);
  else hipLaunchKernelGGL(k_adj_jacobian<3>, blocks(n), dim3(256), 0, s, T, A, jac);
}
void adj_prepare(uint32_t nc, const double* jac, double* minv, hipStream_t s) {
  if (nc) hipLaunchKernelGGL(k_adj_prepare, blocks(nc), dim3(256), 0, s, nc, jac, minv);
}
void adj_init(uint32_t n, const uint8_t* fixed, const double* xbar, const double* vbar, const double* wbar, double dt,
              double* alpha, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_init, blocks(n), dim3(256), 0, s, n, fixed, xbar, vbar, wbar, dt, alpha);
}
void adj_sweep(const AdjTape& T, uint32_t begin, uint32_t end, const double* jac, const double* minv, double* alpha,
               double* lambda, unsigned long long* delta, hipStream_t s) {
  if (end > begin)
    hipLaunchKernelGGL(k_adj_sweep, blocks(end - begin), dim3(256), 0, s, T, begin, end, jac, minv, alpha, lambda,
                       delta);
}
void adj_assemble(const AdjTape& T, const double* jac, const double* alpha, const double* lambda, double* xbar,
                  double* mubar, hipStream_t s) {
  if (T.nc) hipLaunchKernelGGL(k_adj_assemble, blocks(T.nc), dim3(256), 0, s, T, jac, alpha, lambda, xbar, mubar);
}
void adj_finish(uint32_t n, const double* alpha, double* vbar, double* wbar, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_finish, blocks(n), dim3(256), 0, s, n, alpha, vbar, wbar);
}
void adj_restore_R(uint32_t n, const double* q, double* R, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_restore_R, blocks(n), dim3(256), 0, s, n, q, R);
}
void adj_pack_vel(uint32_t n, const double* v, const double* w, double* vel, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_pack_vel, blocks(n), dim3(256), 0, s, n, v, w, vel);
}
void adj_pack_p(uint32_t nc, const double* f, size_t cap, double* p, double* mu, hipStream_t s) {
  if (nc) hipLaunchKernelGGL(k_adj_pack_p, blocks(nc), dim3(256), 0, s, nc, f, cap, p, mu);
}

}  // namespace pe_amd
