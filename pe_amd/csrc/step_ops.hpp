// Per-body and per-contact operations of the time step, shared by the single-world
// kernels and the batched (one workgroup per scene) kernel.  Citations are pe
// source-tree file:line.
#pragma once

#include "kernels.hpp"

namespace pe_amd {

// AABB of body i from its current state.
__device__ __forceinline__ void update_aabb(const Bodies& B, uint32_t i) {
  const m33 R = ld9(B.R, i);
  double o[6];
  if (B.shape[i] == kShapeUnion) union_aabb(B, i, o);
  else body_aabb(B.shape[i], ld3(B.pos, i), R, B.param + 3 * (size_t)i, ld3(B.pnormal, i), B.pd[i], o);
#pragma unroll
  for (int k = 0; k < 6; ++k) B.aabb[6 * (size_t)i + k] = o[k];
}

// Body caching, HCSITS.h:1834-1878 with initializeVelocityCorrections, HCSITS.h:4094-4109.
// Writes the solver body record (predicted v, w; inverse mass; R*Iinv*R^T) and the
// initial velocity corrections from external forces; resets the forces.
__device__ __forceinline__ void predict_body(const Bodies& B, uint32_t i, d3 gravity, double dt) {
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
// loop (HCSITS.h:1954-2004) and integratePositions (HCSITS.h:4135-4184); ends with the
// body's new AABB.
template <class Trig>
__device__ __forceinline__ void integrate_body(const Bodies& B, uint32_t i, double dt, double damping) {
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
  double o[6];
  if (shape == kShapeUnion) {
    // SuperBody::updateBody -> Sphere::update(dq), Sphere.cpp:1052-1070: every member at the
    // union's new position plus its rotated relative position; then the merged box.
    for (uint32_t m = B.mstart[i]; m < B.mstart[i + 1]; ++m) st3(B.mpos, m, gp + mat_vec(R, ld3(B.mrpos, m)));
    union_aabb(B, i, o);
  } else {
    if (shape == kShapePlane) {
      nrm = quat_rotate(qn, mk3(0, 0, 1));
      pd = dot3(nrm, gp);
      st3(B.pnormal, i, nrm);
      B.pd[i] = pd;
    }
    body_aabb(shape, gp, R, B.param + 3 * (size_t)i, nrm, pd, o);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) B.aabb[6 * (size_t)i + k] = o[k];
}

// Contact caching, HCSITS.h:1741-1818: solver slot s receives raw contact c -- lever
// arms, contact frame, error-reduced distance, friction and the Delassus block in the
// contact frame with its inverses.  Impulses start at zero (no warm start).
// counters[0]: max penetration (bits of a non-negative double); counters[1]: contacts
// that involve a fixed body.
__device__ __forceinline__ void setup_contact(const Bodies& B, const RawContacts& R, uint32_t c, size_t s,
                                              const SolverContacts& C, double erp, double dtinv,
                                              unsigned long long* counters) {
  unsigned long long* maxpen = counters;
  const uint32_t i1 = R.b1[c], i2 = R.b2[c];
  if (B.fixed[i1] || B.fixed[i2]) atomicAdd(&counters[1], 1ull);
  const d3 pos = ld3(R.pos, c);
  const d3 r1 = pos - ld3(B.pos, i1);
  const d3 r2 = pos - ld3(B.pos, i2);
  const d3 n = ld3(R.nrm, c);
  d3 t;
  if (fabs(n.x) < fabs(n.y)) {
    if (fabs(n.x) < fabs(n.z)) t = mk3(0, n.z, -n.y);
    else t = mk3(n.y, -n.x, 0);
  } else {
    if (fabs(n.y) < fabs(n.z)) t = mk3(-n.z, 0, n.x);
    else t = mk3(n.y, -n.x, 0);
  }
  t = nrm3(t);
  const d3 o = cross3(n, t);
  const m33 cf = mat_cols(n, t, o);
  double dist = R.dist[c];
  if (dist < 0) {
    atomicMax(maxpen, (unsigned long long)__double_as_longlong(-dist));
    dist *= erp;
  }
  const double* s1 = B.sb + (size_t)kSB * i1;
  const double* s2 = B.sb + (size_t)kSB * i2;
  m33 I1, I2;
#pragma unroll
  for (int k = 0; k < 9; ++k) { I1.a[k] = s1[SB_IW + k]; I2.a[k] = s2[SB_IW + k]; }
  const double im1 = s1[SB_IM], im2 = s2[SB_IM];
  const m33 A = mul_skew(skew_mul(r1, I1), r1);
  const m33 Bm = mul_skew(skew_mul(r2, I2), r2);
  m33 D;
#pragma unroll
  for (int k = 0; k < 9; ++k) D.a[k] = -(A.a[k] + Bm.a[k]);
  D.a[0] += im1 + im2;
  D.a[4] += im1 + im2;
  D.a[8] += im1 + im2;
  D = mat_mul(mat_mul(mat_trans(cf), D), cf);
  const m33 Di = mat_inv(D);
  double det = D.a[4] * D.a[8] - D.a[5] * D.a[7];  // Matrix2x2::getInverse, Matrix2x2.h:977-988
  det = 1.0 / det;
  const size_t cap = C.cap;
  double* f = C.f;
  const d3 vals[5] = {r1, r2, n, t, o};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    f[(3 * k) * cap + s] = vals[k].x;
    f[(3 * k + 1) * cap + s] = vals[k].y;
    f[(3 * k + 2) * cap + s] = vals[k].z;
  }
  f[CF_DDT * cap + s] = dist * dtinv;
  f[CF_MU * cap + s] = R.mu[c];
  f[CF_DNI * cap + s] = 1.0 / D.a[0];
#pragma unroll
  for (int k = 0; k < 9; ++k) f[(CF_DINV + k) * cap + s] = Di.a[k];
  f[(CF_DTOI + 0) * cap + s] = det * D.a[8];
  f[(CF_DTOI + 1) * cap + s] = -det * D.a[5];
  f[(CF_DTOI + 2) * cap + s] = -det * D.a[7];
  f[(CF_DTOI + 3) * cap + s] = det * D.a[4];
  f[(CF_DTO + 0) * cap + s] = D.a[4];
  f[(CF_DTO + 1) * cap + s] = D.a[5];
  f[(CF_DTO + 2) * cap + s] = D.a[7];
  f[(CF_DTO + 3) * cap + s] = D.a[8];
  f[CF_D0 * cap + s] = D.a[0];
  f[(CF_D0 + 1) * cap + s] = D.a[1];
  f[(CF_D0 + 2) * cap + s] = D.a[2];
  f[(CF_D0 + 3) * cap + s] = D.a[3];
  f[(CF_D0 + 4) * cap + s] = D.a[6];
  f[CF_P * cap + s] = 0.0;
  f[(CF_P + 1) * cap + s] = 0.0;
  f[(CF_P + 2) * cap + s] = 0.0;
  C.b1[s] = i1 | (B.fixed[i1] ? kFixedBit : 0u);
  C.b2[s] = i2 | (B.fixed[i2] ? kFixedBit : 0u);
}

}  // namespace pe_amd
