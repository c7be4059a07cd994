// Device-side data layout of one pe world in HBM, and the per-pair / per-contact
// routines shared by the kernels.  Citations are pe source-tree file:line.
#pragma once

#include <stdint.h>

#include "dmath.hpp"
#include "trig_dd.hpp"

namespace pe_amd {

constexpr double kContactThreshold = 1e-8;   // pe/core/Thresholds.h:151
constexpr double kFrictionThreshold = 1e-8;  // pe/core/Thresholds.h:161
constexpr uint32_t kFixedBit = 0x80000000u;  // set in solver body indices of fixed bodies
constexpr int kShapeSphere = 1, kShapeBox = 2, kShapePlane = 3, kShapeCapsule = 4, kShapeUnion = 5;
constexpr double kParallelThreshold = 1e-8;  // pe/core/Thresholds.h:170

// Solver body record (128 B, one L2 line): predicted v, w; inverse mass; world-frame
// inverse inertia R*Iinv*R^T as 9 doubles (pe recomputes the full product on every
// access, RigidBody.h:866-869, so it is not symmetric to the last bit).
constexpr int kSB = 16;
constexpr int SB_V = 0, SB_W = 3, SB_IM = 6, SB_IW = 7;
// Velocity-correction record (64 B): dv, dw, 2 pad.
constexpr int kDVW = 8;

// Bodies: per-body arrays in the caller's body order.
struct Bodies {
  uint32_t n;
  const uint8_t* shape;
  const uint8_t* fixed;
  const uint32_t* mat;
  const uint64_t* sid;
  const double* param;     // 3n: radius | side lengths
  double* pos;             // 3n
  double* q;               // 4n
  double* R;               // 9n
  double* v;               // 3n
  double* w;               // 3n
  double* pnormal;         // 3n (planes)
  double* pd;              // n  (planes)
  double* aabb;            // 6n
  const double* inv_mass;  // n
  const double* Ib;        // 9n body-frame inertia
  const double* Ibinv;     // 9n body-frame inverse inertia
  double* force;           // 3n
  double* torque;          // 3n
  double* sb;              // kSB * n
  double* dvw;             // kDVW * n
  // Union members (spheres), grouped by union in body order: the members of body i are
  // [mstart[i], mstart[i+1]) (empty for every other shape).  Geometry id n + k names member k.
  uint32_t nm = 0;
  const uint32_t* mstart = nullptr;  // n + 1 (null when nm == 0)
  const uint32_t* mowner = nullptr;  // nm: the union's body index
  const uint64_t* msid = nullptr;    // nm
  const uint32_t* mmat = nullptr;    // nm
  const double* mrad = nullptr;      // nm
  const double* mrpos = nullptr;     // 3nm: position in the union's body frame
  double* mpos = nullptr;            // 3nm: global position
};

PE_HD d3 ld3(const double* p, uint32_t i) { return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
PE_HD void st3(double* p, uint32_t i, d3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
PE_HD m33 ld9(const double* p, uint32_t i) {
  m33 m;
#pragma unroll
  for (int k = 0; k < 9; ++k) m.a[k] = p[9 * (size_t)i + k];
  return m;
}
PE_HD qd ld4(const double* p, uint32_t i) {
  qd q;
#pragma unroll
  for (int k = 0; k < 4; ++k) q.v[k] = p[4 * (size_t)i + k];
  return q;
}

// Closed-interval AABB overlap (BoundingBox::overlaps, BoundingBox.h:487-497).
PE_HD bool aabb_overlap(const double* a, const double* b) {
  return !(b[0] > a[3] || b[1] > a[4] || b[2] > a[5] || b[3] < a[0] || b[4] < a[1] || b[5] < a[2]);
}

// Largest AABB extent (BodyTrait<HashGrids>::getAABBSize, coarse/BodyTrait.h:209-218).
PE_HD double aabb_size(const double* a) {
  double s = a[3] - a[0];
  const double y = a[4] - a[1], z = a[5] - a[2];
  if (y > s) s = y;
  if (z > s) s = z;
  return s;
}

// A sphere taking part in a contact: a sphere body or a union member.  g is its geometry id
// (the body index, or n + member index); contacts join the top-level bodies (ContactBase.h:209-210).
struct SphereRef {
  d3 x;
  double r;
  uint64_t sid;
  uint32_t g;
};
PE_HD SphereRef body_sphere(const Bodies& B, uint32_t i) {
  return SphereRef{ld3(B.pos, i), B.param[3 * (size_t)i], B.sid[i], i};
}
PE_HD SphereRef member_sphere(const Bodies& B, uint32_t m) {
  return SphereRef{ld3(B.mpos, m), B.mrad[m], B.msid[m], B.n + m};
}
PE_HD uint32_t geom_body(const Bodies& B, uint32_t g) { return g < B.n ? g : B.mowner[g - B.n]; }
PE_HD uint32_t geom_mat(const Bodies& B, uint32_t g) { return g < B.n ? B.mat[g] : B.mmat[g - B.n]; }

// Union::calcBoundingBox (UnionBase.cpp:208-230): the members' sphere boxes (SphereBase.h:208-221)
// merged in member order (BoundingBox::operator+=, BoundingBox.h:339-348).
PE_HD void union_aabb(const Bodies& B, uint32_t i, double* o) {
  const uint32_t m0 = B.mstart[i], m1 = B.mstart[i + 1];
  for (uint32_t m = m0; m < m1; ++m) {
    const d3 x = ld3(B.mpos, m);
    const double l = B.mrad[m] + kContactThreshold;
    const double a[6] = {x.x - l, x.y - l, x.z - l, x.x + l, x.y + l, x.z + l};
    if (m == m0) {
      for (int k = 0; k < 6; ++k) o[k] = a[k];
    } else {
      for (int k = 0; k < 3; ++k) o[k] = (a[k] < o[k]) ? a[k] : o[k];
      for (int k = 3; k < 6; ++k) o[k] = (o[k] < a[k]) ? a[k] : o[k];
    }
  }
}

// AABB of one body (SphereBase.h:208-221, BoxBase.cpp:127-141, PlaneBase.cpp:132-148).
PE_HD void body_aabb(int shape, d3 gp, const m33& R, const double* prm, d3 nrm, double d, double* o) {
  if (shape == kShapeSphere) {
    const double l = prm[0] + kContactThreshold;
    o[0] = gp.x - l; o[1] = gp.y - l; o[2] = gp.z - l;
    o[3] = gp.x + l; o[4] = gp.y + l; o[5] = gp.z + l;
  } else if (shape == kShapeBox) {
    const double* r = R.a;
    const double xl = 0.5 * (fabs(r[0] * prm[0]) + fabs(r[1] * prm[1]) + fabs(r[2] * prm[2])) + kContactThreshold;
    const double yl = 0.5 * (fabs(r[3] * prm[0]) + fabs(r[4] * prm[1]) + fabs(r[5] * prm[2])) + kContactThreshold;
    const double zl = 0.5 * (fabs(r[6] * prm[0]) + fabs(r[7] * prm[1]) + fabs(r[8] * prm[2])) + kContactThreshold;
    o[0] = gp.x - xl; o[1] = gp.y - yl; o[2] = gp.z - zl;
    o[3] = gp.x + xl; o[4] = gp.y + yl; o[5] = gp.z + zl;
  } else if (shape == kShapeCapsule) {  // CapsuleBase.cpp:124-138 (axis = body x-axis)
    const double* r = R.a;
    const double xl = fabs(r[0] * prm[1]) * 0.5 + prm[0] + kContactThreshold;
    const double yl = fabs(r[3] * prm[1]) * 0.5 + prm[0] + kContactThreshold;
    const double zl = fabs(r[6] * prm[1]) * 0.5 + prm[0] + kContactThreshold;
    o[0] = gp.x - xl; o[1] = gp.y - yl; o[2] = gp.z - zl;
    o[3] = gp.x + xl; o[4] = gp.y + yl; o[5] = gp.z + zl;
  } else {
    const double inf = __builtin_huge_val();
    o[0] = (nrm.x < 0 && nrm.y == 0 && nrm.z == 0) ? (-d - kContactThreshold) : -inf;
    o[1] = (nrm.x == 0 && nrm.y < 0 && nrm.z == 0) ? (-d - kContactThreshold) : -inf;
    o[2] = (nrm.x == 0 && nrm.y == 0 && nrm.z < 0) ? (-d - kContactThreshold) : -inf;
    o[3] = (nrm.x > 0 && nrm.y == 0 && nrm.z == 0) ? (d + kContactThreshold) : inf;
    o[4] = (nrm.x == 0 && nrm.y > 0 && nrm.z == 0) ? (d + kContactThreshold) : inf;
    o[5] = (nrm.x == 0 && nrm.y == 0 && nrm.z > 0) ? (d + kContactThreshold) : inf;
  }
}

// ---------------------------------------------------------------------------------------------
// Narrowphase (pe/core/detection/fine/MaxContacts.h).  Each routine calls
// emit(b1, b2, gpos, normal, dist) for every contact it creates, in pe's order.
// ---------------------------------------------------------------------------------------------

// collideSphereSphere, MaxContacts.h:547-568: body1 is the one with the smaller SID.
template <class Emit>
__device__ void np_sphere_sphere(SphereRef s1, SphereRef s2, Emit& emit) {
  if (s2.sid < s1.sid) { const SphereRef t = s1; s1 = s2; s2 = t; }
  d3 normal = s1.x - s2.x;
  const double dist = len3(normal) - s1.r - s2.r;
  if (dist < kContactThreshold) {
    normal = nrm3(normal);
    const double k = s2.r + 0.5 * dist;
    emit(s1.g, s2.g, s2.x + normal * k, normal, dist);
  }
}

// collideSphereBox, MaxContacts.h:589-667: the sphere is body1.
template <class Emit>
__device__ void np_sphere_box(const Bodies& B, const SphereRef& S, uint32_t b, Emit& emit) {
  const double* L = B.param + 3 * (size_t)b;
  const double l0 = 0.5 * L[0], l1 = 0.5 * L[1], l2 = 0.5 * L[2];
  const d3 spos = S.x, bpos = ld3(B.pos, b);
  const uint32_t s = S.g;
  const m33 R = ld9(B.R, b);
  const double* r = R.a;
  bool outside = false;
  const d3 d = spos - bpos;
  double p[3] = {d.x * r[0] + d.y * r[3] + d.z * r[6], d.x * r[1] + d.y * r[4] + d.z * r[7],
                 d.x * r[2] + d.y * r[5] + d.z * r[8]};
  const double l[3] = {l0, l1, l2};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (p[i] < -l[i]) { p[i] = -l[i]; outside = true; }
    else if (p[i] > l[i]) { p[i] = l[i]; outside = true; }
  }
  const double radius = S.r;
  if (!outside) {
    double dist = fabs(p[0]) - l[0];
    int face = 0;
    for (int i = 1; i < 3; ++i) {
      const double tmp = fabs(p[i]) - l[i];
      if (dist < tmp) { dist = tmp; face = i; }
    }
    double nn[3] = {0.0, 0.0, 0.0};
    nn[face] = (p[face] > 0.0) ? 1.0 : -1.0;
    const d3 n = mat_vec(R, mk3(nn[0], nn[1], nn[2]));
    dist -= radius;
    emit(s, b, spos, n, dist);
    return;
  }
  const d3 qv = mat_vec(R, mk3(p[0], p[1], p[2]));
  const d3 n = d - qv;
  const double dist = len3(n) - radius;
  if (dist < kContactThreshold) emit(s, b, bpos + qv, nrm3(n), dist);
}

// collideSpherePlane, MaxContacts.h:777-792: the sphere is body1.
template <class Emit>
__device__ void np_sphere_plane(const Bodies& B, const SphereRef& S, uint32_t p, Emit& emit) {
  const d3 n = ld3(B.pnormal, p);
  const d3 x = S.x;
  const double r = S.r;
  const double k = dot3(n, x);
  const double dist = k - r - B.pd[p];
  if (dist < kContactThreshold) emit(S.g, p, x - (r + dist) * n, n, dist);
}

// collideBoxPlane, MaxContacts.h:2380-2528: corners in pe's order, each pre-filtered.
template <class Emit>
__device__ void np_box_plane(const Bodies& B, uint32_t b, uint32_t p, Emit& emit) {
  const d3 bpos = ld3(B.pos, b);
  const double* L = B.param + 3 * (size_t)b;
  const m33 R = ld9(B.R, b);
  const double xr = 0.5 * L[0], yr = 0.5 * L[1], zr = 0.5 * L[2];
  const d3 pn = ld3(B.pnormal, p);
  const double pdist = B.pd[p];
  const d3 nrel = mat_tvec(R, pn);
  const double mn = fmin(fmin(xr, yr), zr);
  const double minlength = -0.99 * mn;
  // corner sign patterns (-,-,-) (+,-,-) (+,+,-) (-,+,-) (-,-,+) (+,-,+) (+,+,+) (-,+,+)
  const unsigned sx = 0x66u, sy = 0xCCu, sz = 0xF0u;  // bit c set = '+' for corner c
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const bool px = (sx >> c) & 1u, py = (sy >> c) & 1u, pz = (sz >> c) & 1u;
    const double cx = px ? xr : -xr, cy = py ? yr : -yr, cz = pz ? zr : -zr;
    double t = cx * nrel.x;
    t = py ? t + yr * nrel.y : t - yr * nrel.y;
    t = pz ? t + zr * nrel.z : t - zr * nrel.z;
    if (t < minlength) {
      const d3 pos = bpos + mat_vec(R, mk3(cx, cy, cz));
      const double k = dot3(pos, pn);
      const double dist = k - pdist;
      if (dist < kContactThreshold) emit(b, p, pos, pn, dist);
    }
  }
}

// Row-major 3x3 * 3-vector and its transpose on plain arrays.
PE_HD void mv3(const double* m, const double* v, double* o) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}
PE_HD void mtv3(const double* m, const double* v, double* o) {
  o[0] = m[0] * v[0] + m[3] * v[1] + m[6] * v[2];
  o[1] = m[1] * v[0] + m[4] * v[1] + m[7] * v[2];
  o[2] = m[2] * v[0] + m[5] * v[1] + m[8] * v[2];
}

// collideBoxBox, MaxContacts.h:867-2268: separating axes (3 + 3 face axes, 9 edge axes), then
// one edge-edge contact (at the closest points of the two edges, intersectLines,
// src/core/GeomTools.cpp:387-408) or the clipped face contacts: the incident face's edges
// against the reference face's sides, incident vertices inside the reference face, and
// reference corners inside the incident face.
template <class Emit>
__device__ void np_box_box(const Bodies& B, uint32_t i1, uint32_t i2, Emit& emit) {
  if (B.sid[i2] < B.sid[i1]) { uint32_t t = i1; i1 = i2; i2 = t; }
  double R1[9], R2[9], R[9], Q[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) { R1[k] = B.R[9 * (size_t)i1 + k]; R2[k] = B.R[9 * (size_t)i2 + k]; }
  {
    m33 a, b;
    for (int k = 0; k < 9; ++k) { a.a[k] = R1[k]; b.a[k] = R2[k]; }
    const m33 r = mat_mul(mat_trans(a), b);
    for (int k = 0; k < 9; ++k) { R[k] = r.a[k]; Q[k] = fabs(r.a[k]); }
  }
  const double* p1 = B.pos + 3 * (size_t)i1;
  const double* p2 = B.pos + 3 * (size_t)i2;
  double d12[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  double P[3];
  mtv3(R1, d12, P);
  const double* L1 = B.param + 3 * (size_t)i1;
  const double* L2 = B.param + 3 * (size_t)i2;
  const double hl1[3] = {0.5 * L1[0], 0.5 * L1[1], 0.5 * L1[2]};
  const double hl2[3] = {0.5 * L2[0], 0.5 * L2[1], 0.5 * L2[2]};
  double maxDepth = -__builtin_huge_val();
  bool invertNormal = false;
  unsigned contactCase = 0;
  double cn[3] = {0, 0, 0}, nc[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {  // face axes of box 1
    const double term1 = fabs(P[a]) - (hl1[a] + hl2[0] * Q[3 * a] + hl2[1] * Q[3 * a + 1] + hl2[2] * Q[3 * a + 2]);
    if (term1 > kContactThreshold) return;
    if (a == 0 || term1 > maxDepth) {
      maxDepth = term1;
      cn[0] = R1[a]; cn[1] = R1[3 + a]; cn[2] = R1[6 + a];
      contactCase = 1 + a;
      invertNormal = (P[a] < 0.0);
    }
  }
  for (int a = 0; a < 3; ++a) {  // face axes of box 2
    const double term1 = P[0] * R[a] + P[1] * R[3 + a] + P[2] * R[6 + a];
    const double term2 = fabs(term1) - (hl1[0] * Q[a] + hl1[1] * Q[3 + a] + hl1[2] * Q[6 + a] + hl2[a]);
    if (term2 > kContactThreshold) return;
    if (term2 > maxDepth) {
      maxDepth = term2;
      cn[0] = R2[a]; cn[1] = R2[3 + a]; cn[2] = R2[6 + a];
      contactCase = 4 + a;
      invertNormal = (term1 < 0.0);
    }
  }
  // edge axes (cases 7-15), table of index tuples (see the column meaning below)
  //   term1 = P[0]*R[1] - P[2]*R[3];  term2 = hl1[4]*Q[5] + hl1[6]*Q[7];
  //   term3 = hl2[8]*Q[9] + hl2[10]*Q[11];  length = sqrt(R[12]^2 + R[13]^2)
  constexpr int8_t E[9][14] = {
      {2, 3, 1, 6, 1, 6, 2, 3, 2, 1, 1, 2, 6, 3}, {2, 4, 1, 7, 1, 7, 2, 4, 0, 2, 2, 0, 7, 4},
      {2, 5, 1, 8, 1, 8, 2, 5, 1, 0, 0, 1, 8, 5}, {0, 6, 2, 0, 2, 0, 0, 6, 2, 4, 1, 5, 6, 0},
      {0, 7, 2, 1, 2, 1, 0, 7, 0, 5, 2, 3, 7, 1}, {0, 8, 2, 2, 2, 2, 0, 8, 1, 3, 0, 4, 8, 2},
      {1, 0, 0, 3, 0, 3, 1, 0, 2, 7, 1, 8, 3, 0}, {1, 1, 0, 4, 0, 4, 1, 1, 0, 8, 2, 6, 4, 1},
      {1, 2, 0, 5, 0, 5, 1, 2, 1, 6, 0, 7, 5, 2}};
  // normal_c per case: (sign, index of R) per component, sign 0 = zero component
  constexpr int8_t NC[9][6] = {{0, 0, -1, 6, 1, 3}, {0, 0, -1, 7, 1, 4}, {0, 0, -1, 8, 1, 5},
                               {1, 6, 0, 0, -1, 0}, {1, 7, 0, 0, -1, 1}, {1, 8, 0, 0, -1, 2},
                               {-1, 3, 1, 0, 0, 0}, {-1, 4, 1, 1, 0, 0}, {-1, 5, 1, 2, 0, 0}};
  for (int e = 0; e < 9; ++e) {
    const int8_t* t = E[e];
    const double term1 = P[t[0]] * R[t[1]] - P[t[2]] * R[t[3]];
    const double term2 = hl1[t[4]] * Q[t[5]] + hl1[t[6]] * Q[t[7]];
    const double term3 = hl2[t[8]] * Q[t[9]] + hl2[t[10]] * Q[t[11]];
    double sum = fabs(term1) - (term2 + term3);
    if (fabs(term1) > term2 + term3 + kContactThreshold) return;
    const double length = sqrt(R[t[12]] * R[t[12]] + R[t[13]] * R[t[13]]);
    if (length > 2.220446049250313e-16) {  // machine epsilon (pe's `epsilon`)
      sum /= length;
      if (sum > maxDepth && fabs(sum - maxDepth) > 1e-8) {  // pe's `accuracy`
        maxDepth = sum;
        for (int k = 0; k < 3; ++k) {
          const int sg = NC[e][2 * k], ix = NC[e][2 * k + 1];
          nc[k] = sg == 0 ? 0.0 : (sg > 0 ? R[ix] / length : -R[ix] / length);
        }
        contactCase = 7 + e;
        invertNormal = (term1 < 0.0);
      }
    }
  }
  if (contactCase == 0) return;
  if (contactCase > 6) mv3(R1, nc, cn);
  if (invertNormal) { cn[0] = -cn[0]; cn[1] = -cn[1]; cn[2] = -cn[2]; }
  const d3 ncontact = mk3(cn[0], cn[1], cn[2]);

  if (contactCase > 6) {  // edge-edge
    double sg[3], off[3];
    mtv3(R1, cn, sg);
    for (int i = 0; i < 3; ++i) sg[i] = (sg[i] > 0) ? hl1[i] : -hl1[i];
    mv3(R1, sg, off);
    d3 pB1 = mk3(p1[0], p1[1], p1[2]) + mk3(off[0], off[1], off[2]);
    mtv3(R2, cn, sg);
    for (int i = 0; i < 3; ++i) sg[i] = (sg[i] > 0) ? -hl2[i] : hl2[i];
    mv3(R2, sg, off);
    d3 pB2 = mk3(p2[0], p2[1], p2[2]) + mk3(off[0], off[1], off[2]);
    const d3 ua = mk3(R1[(contactCase - 7) / 3], R1[(contactCase - 7) / 3 + 3], R1[(contactCase - 7) / 3 + 6]);
    const d3 ub = mk3(R2[(contactCase - 7) % 3], R2[(contactCase - 7) % 3 + 3], R2[(contactCase - 7) % 3 + 6]);
    double s = 0.0, t = 0.0;
    const d3 cr = cross3(ua, ub);
    const double sqrlen = dot3(cr, cr);
    if (!(sqrlen < 1e-8)) {  // parallelThreshold
      const double isqrlen = 1.0 / sqrlen;
      const d3 pp = pB2 - pB1;
      const double dt = dot3(ua, ub);
      const double a = dot3(ua, pp);
      const double b = -dot3(ub, pp);
      s = (a * dot3(ub, ub) + dt * b) * isqrlen;
      t = (b * dot3(ua, ua) + dt * a) * isqrlen;
    }
    pB1 = pB1 + s * ua;
    pB2 = pB2 + t * ub;
    emit(i2, i1, 0.5 * (pB1 + pB2), ncontact, maxDepth);
    return;
  }

  // face case: reference face on box "a", incident face on box "b"
  const double* hla = hl1;
  const double* hlb = hl2;
  const double* Ra = R1;
  const double* Rb = R2;
  const double* pa = p1;
  const double* pb = p2;
  double n[3] = {cn[0], cn[1], cn[2]};
  if (contactCase > 3) {
    uint32_t tt = i1; i1 = i2; i2 = tt;
    hla = hl2; hlb = hl1; Ra = R2; Rb = R1; pa = p2; pb = p1;
    n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
  }
  const d3 nn = mk3(n[0], n[1], n[2]);
  double na[3], nb[3];
  mtv3(Ra, n, na);
  mtv3(Rb, n, nb);
  unsigned xb, yb, zb;
  if (fabs(nb[0]) > fabs(nb[1])) {
    if (fabs(nb[0]) > fabs(nb[2])) { xb = 0; yb = 1; zb = 2; }
    else { xb = 2; yb = 0; zb = 1; }
  } else if (fabs(nb[1]) > fabs(nb[2])) {
    xb = 1; yb = 2; zb = 0;
  } else {
    xb = 2; yb = 0; zb = 1;
  }
  const unsigned xa = (contactCase < 4) ? contactCase - 1 : contactCase - 4;
  const unsigned ya = (xa == 0) ? 1 : (xa == 1 ? 2 : 0);
  const unsigned za = (xa == 0) ? 2 : (xa == 1 ? 0 : 1);
  double ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
  double vba[4][3];
  for (int k = 0; k < 4; ++k) {
    double vb[3], w[3], u[3];
    vb[xb] = (nb[xb] > 0.0) ? -hlb[xb] : hlb[xb];
    vb[yb] = (k & 1) ? hlb[yb] : -hlb[yb];
    vb[zb] = (k & 2) ? hlb[zb] : -hlb[zb];
    mv3(Rb, vb, w);
    u[0] = ab[0] + w[0]; u[1] = ab[1] + w[1]; u[2] = ab[2] + w[2];
    mtv3(Ra, u, vba[k]);
  }
  const double offseta = (na[xa] > 0.0) ? hla[xa] : -hla[xa];
  constexpr int8_t edges[4][2] = {{0, 1}, {0, 2}, {3, 1}, {3, 2}};
  for (int e = 0; e < 4; ++e) {
    const double* v0 = vba[edges[e][0]];
    const double* v1 = vba[edges[e][1]];
    for (int side = 0; side < 4; ++side) {
      const unsigned d = (side < 2) ? ya : za;
      const unsigned o = (side < 2) ? za : ya;
      const double bound = (side & 1) ? -hla[d] : hla[d];
      if ((v0[d] > bound) ^ (v1[d] > bound)) {
        const double s = (bound - v0[d]) / (v1[d] - v0[d]);
        if (s > 0.0 && s < 1.0) {
          const double tmp = v0[o] + s * (v1[o] - v0[o]);
          if (fabs(tmp) < hla[o]) {
            const double dist = fabs(v0[xa] + s * (v1[xa] - v0[xa])) - hla[xa];
            if (dist < kContactThreshold) {
              double posa[3], w[3];
              posa[xa] = offseta;
              posa[d] = bound;
              posa[o] = tmp;
              mv3(Ra, posa, w);
              emit(i2, i1, mk3(pa[0], pa[1], pa[2]) + mk3(w[0], w[1], w[2]), nn, dist);
            }
          }
        }
      }
    }
  }
  for (int k = 0; k < 4; ++k) {  // incident vertices inside the reference face
    if (fabs(vba[k][ya]) <= hla[ya] && fabs(vba[k][za]) <= hla[za]) {
      const double dist = fabs(vba[k][xa]) - hla[xa];
      if (dist < kContactThreshold) {
        double w[3];
        mv3(Ra, vba[k], w);
        emit(i2, i1, mk3(pa[0], pa[1], pa[2]) + mk3(w[0], w[1], w[2]), nn, dist);
      }
    }
  }
  double ba[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
  for (int k = 0; k < 4; ++k) {  // reference corners inside the incident face
    double va[3], w[3], u[3], vab[3];
    va[xa] = (na[xa] > 0.0) ? hla[xa] : -hla[xa];
    va[ya] = (k & 1) ? hla[ya] : -hla[ya];
    va[za] = (k & 2) ? hla[za] : -hla[za];
    mv3(Ra, va, w);
    u[0] = ba[0] + w[0]; u[1] = ba[1] + w[1]; u[2] = ba[2] + w[2];
    mtv3(Rb, u, vab);
    if (((fabs(vab[yb]) <= hlb[yb] && fabs(vab[zb]) < hlb[zb]) ||
         (fabs(vab[yb]) < hlb[yb] && fabs(vab[zb]) <= hlb[zb]))) {
      const double dist = fabs(vab[xb]) - hlb[xb];
      if (dist < kContactThreshold) {
        double w2[3];
        mv3(Rb, vab, w2);
        emit(i2, i1, mk3(pb[0], pb[1], pb[2]) + mk3(w2[0], w2[1], w2[2]), nn, dist);
      }
    }
  }
}

// MaxContacts::collide dispatch for (a, b) as handed over by the broadphase
// (MaxContacts.h:242-526).
// getClosestLineSegmentPoints, src/core/GeomTools.cpp:259-385: closest points of the segments
// [a1, a2] and [b1, b2] (end-point cases, end point against interior, then both interiors).
__device__ inline void closest_segment_points(d3 a1, d3 a2, d3 b1, d3 b2, d3& cp1, d3& cp2) {
  const d3 a1a2 = a2 - a1, b1b2 = b2 - b1, a1b1 = b1 - a1;
  const double da1 = dot3(a1a2, a1b1), db1 = dot3(b1b2, a1b1);
  if (da1 <= 0.0 && db1 >= 0.0) { cp1 = a1; cp2 = b1; return; }
  const d3 a1b2 = b2 - a1;
  const double da2 = dot3(a1a2, a1b2), db2 = dot3(b1b2, a1b2);
  if (da2 <= 0.0 && db2 <= 0.0) { cp1 = a1; cp2 = b2; return; }
  const d3 a2b1 = b1 - a2;
  const double da3 = dot3(a1a2, a2b1), db3 = dot3(b1b2, a2b1);
  if (da3 >= 0.0 && db3 >= 0.0) { cp1 = a2; cp2 = b1; return; }
  const d3 a2b2 = b2 - a2;
  const double da4 = dot3(a1a2, a2b2), db4 = dot3(b1b2, a2b2);
  if (da4 >= 0.0 && db4 <= 0.0) { cp1 = a2; cp2 = b2; return; }
  const double la = dot3(a1a2, a1a2);
  if (da1 >= 0.0 && da3 <= 0.0) {
    const d3 k = (da1 / la) * a1a2;
    if (dot3(b1b2, a1b1 - k) >= 0.0) { cp1 = a1 + k; cp2 = b1; return; }
  }
  if (da2 >= 0.0 && da4 <= 0.0) {
    const d3 k = (da2 / la) * a1a2;
    if (dot3(b1b2, a1b2 - k) <= 0.0) { cp1 = a1 + k; cp2 = b2; return; }
  }
  const double lb = dot3(b1b2, b1b2);
  if (db1 <= 0.0 && db2 >= 0.0) {
    const d3 k = (-db1 / lb) * b1b2;
    if (dot3(a1a2, -a1b1 - k) >= 0.0) { cp1 = a1; cp2 = b1 + k; return; }
  }
  if (db3 <= 0.0 && db4 >= 0.0) {
    const d3 k = (-db3 / lb) * b1b2;
    if (dot3(a1a2, -a2b1 - k) <= 0.0) { cp1 = a2; cp2 = b1 + k; return; }
  }
  const double scale = dot3(a1a2, b1b2);
  double div = la * lb - scale * scale;
  div = 1.0 / div;
  const double sa = (lb * da1 - scale * db1) * div;
  const double tb = (scale * da1 - la * db1) * div;
  cp1 = a1 + sa * a1a2;
  cp2 = b1 + tb * b1b2;
}

// getClosestLineBoxPoints, src/core/GeomTools.cpp:73-257: closest points of the segment [p1, p2]
// and a box (centre c, rotation R, sides).  In the box frame, the segment's direction mirrored
// into the positive octant, the squared distance is piecewise quadratic in the parameter t with
// breaks where the point crosses a face plane; the walk follows the pieces from t = 0 until the
// derivative changes sign, then clamps the segment point to the box.
__device__ inline void closest_line_box_points(d3 p1, d3 p2, d3 c, const m33& R, const double* side, d3& lret,
                                               d3& bret) {
  const double eps = 2.220446049250313e-16;  // pe's `epsilon` (machine epsilon)
  const d3 l = p2 - p1;
  const d3 tl = mat_tvec(R, l);
  const double tv[3] = {tl.x, tl.y, tl.z};
  double sgn[3], v[3], v2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    sgn[i] = tv[i] < 0 ? -1.0 : 1.0;
    v[i] = sgn[i] * tv[i];
    v2[i] = v[i] * v[i];
  }
  const d3 u = p1 - c;
  const double* r = R.a;
  const double sv[3] = {sgn[0] * (r[0] * u.x + r[3] * u.y + r[6] * u.z), sgn[1] * (r[1] * u.x + r[4] * u.y + r[7] * u.z),
                        sgn[2] * (r[2] * u.x + r[5] * u.y + r[8] * u.z)};
  const double h[3] = {0.5 * side[0], 0.5 * side[1], 0.5 * side[2]};
  int region[3] = {0, 0, 0};
  double tanchor[3] = {2, 2, 2};
  double dd2dt = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (v[i] > eps) {
      if (sv[i] < -h[i]) {
        region[i] = -1;
        tanchor[i] = (-h[i] - sv[i]) / v[i];
        dd2dt -= v2[i] * tanchor[i];
      } else if (sv[i] > h[i]) {
        region[i] = 1;
        tanchor[i] = (h[i] - sv[i]) / v[i];
        dd2dt -= v2[i] * tanchor[i];
      } else {
        tanchor[i] = (h[i] - sv[i]) / v[i];
      }
    }
  }
  double t = 0;
  if (dd2dt < -1e-8) {   // pe's `accuracy`
    // at most one region step per axis and direction: the walk ends within 7 pieces
    for (int it = 0; it < 8; ++it) {
      double next_t = 1;
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (tanchor[i] > t && tanchor[i] < 1.0 && tanchor[i] < next_t) next_t = tanchor[i];
      double next_dd2dt = 0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (region[i] != 0) next_dd2dt += v2[i] * (next_t - tanchor[i]);
      if (next_dd2dt >= 0) {
        t -= dd2dt / ((next_dd2dt - dd2dt) / (next_t - t));
        break;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (tanchor[i] == next_t) {
          tanchor[i] = (h[i] - sv[i]) / v[i];
          ++region[i];
        }
      t = next_t;
      dd2dt = next_dd2dt;
      if (!(t < 1.0)) break;
    }
  }
  lret = p1 + t * l;
  double bx[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double x = sgn[i] * (sv[i] + t * v[i]);
    if (x < -h[i]) x = -h[i];
    else if (x > h[i]) x = h[i];
    bx[i] = x;
  }
  bret = mat_vec(R, mk3(bx[0], bx[1], bx[2])) + c;
}

// Vector3 operator== (Vector3.h:1421-1428): every component within 1e-8 (Equal.h:172-175).
__device__ inline bool nearly_equal3(d3 a, d3 b) {
  return fabs(a.x - b.x) < 1e-8 && fabs(a.y - b.y) < 1e-8 && fabs(a.z - b.z) < 1e-8;
}

// collideSphereCapsule, MaxContacts.h:690-725: the sphere (body1) against the closest point of
// the capsule's axis.
template <class Emit>
__device__ void np_sphere_capsule(const Bodies& B, const SphereRef& S, uint32_t c, Emit& emit) {
  const double hl = 0.5 * B.param[3 * (size_t)c + 1];
  const double rc = B.param[3 * (size_t)c];
  const d3 spos = S.x, cpos = ld3(B.pos, c);
  const double* R = B.R + 9 * (size_t)c;
  double sx = R[0] * (spos.x - cpos.x) + R[3] * (spos.y - cpos.y) + R[6] * (spos.z - cpos.z);
  if (sx > hl) sx = hl;
  else if (sx < -hl) sx = -hl;
  const d3 spos2 = mk3(sx * R[0] + cpos.x, sx * R[3] + cpos.y, sx * R[6] + cpos.z);
  d3 normal = spos - spos2;
  const double dist = len3(normal) - S.r - rc;
  if (dist < kContactThreshold) {
    normal = nrm3(normal);
    const double k = rc + 0.5 * dist;
    emit(S.g, c, spos2 + normal * k, normal, dist);
  }
}

// collideBoxCapsule, MaxContacts.h:2273-2330: closest points of the axis and the box from each
// end of the axis; the second contact needs only points that differ from the first's (no
// distance test, as in the reference).  The box is body1.
template <class Emit>
__device__ void np_box_capsule(const Bodies& B, uint32_t b, uint32_t c, Emit& emit) {
  const double* Rc = B.R + 9 * (size_t)c;
  const double rc = B.param[3 * (size_t)c], len = B.param[3 * (size_t)c + 1];
  const d3 c_up = mk3(0.5 * len * Rc[0], 0.5 * len * Rc[3], 0.5 * len * Rc[6]);
  const d3 cpos = ld3(B.pos, c);
  const d3 up = cpos + c_up, down = cpos - c_up;
  const d3 bpos = ld3(B.pos, b);
  const m33 Rb = ld9(B.R, b);
  const double* side = B.param + 3 * (size_t)b;
  d3 cp1, bp1;
  closest_line_box_points(up, down, bpos, Rb, side, cp1, bp1);
  d3 normal1 = bp1 - cp1;
  const double dist1 = len3(normal1) - rc;
  if (dist1 > kContactThreshold) return;
  normal1 = nrm3(normal1);
  emit(b, c, bp1 - 0.5 * dist1 * normal1, normal1, dist1);
  d3 cp2, bp2;
  closest_line_box_points(down, up, bpos, Rb, side, cp2, bp2);
  if (nearly_equal3(cp1, cp2)) return;
  d3 normal2 = bp2 - cp2;
  const double dist2 = len3(normal2) - rc;
  normal2 = nrm3(normal2);
  emit(b, c, bp2 - 0.5 * dist2 * normal2, normal2, dist2);
}

// collideCapsuleCapsule, MaxContacts.h:2603-2690: body1 has the smaller SID.  Nearly parallel,
// overlapping axes give two contacts at the ends of the overlap, otherwise one contact between
// the axes' closest points.
template <class Emit>
__device__ void np_capsule_capsule(const Bodies& B, uint32_t c1, uint32_t c2, Emit& emit) {
  if (B.sid[c2] < B.sid[c1]) { uint32_t t = c1; c1 = c2; c2 = t; }
  const double* R1 = B.R + 9 * (size_t)c1;
  const double* R2 = B.R + 9 * (size_t)c2;
  const double r1 = B.param[3 * (size_t)c1], L1 = B.param[3 * (size_t)c1 + 1];
  const double r2 = B.param[3 * (size_t)c2], L2 = B.param[3 * (size_t)c2 + 1];
  const d3 d1 = mk3(R1[0], R1[3], R1[6]), d2 = mk3(R2[0], R2[3], R2[6]);
  const d3 up1 = (0.5 * L1) * d1, up2 = (0.5 * L2) * d2;
  const d3 x1 = ld3(B.pos, c1), x2 = ld3(B.pos, c2);
  d3 cp1, cp2;
  closest_segment_points(x1 + up1, x1 - up1, x2 + up2, x2 - up2, cp1, cp2);
  d3 normal = cp1 - cp2;
  const double dist = len3(normal) - r1 - r2;
  if (!(dist < kContactThreshold)) return;
  normal = nrm3(normal);
  const double c2x = dot3(d1, x2 - x1);
  if ((fabs(dot3(d1, d2)) - 1.0) > -kParallelThreshold && c2x > -L1 && c2x < L1) {
    const double k = r1 + 0.5 * dist;
    const double hl1 = 0.5 * L1, hl2 = 0.5 * L2;
    if (hl1 < c2x + hl2) emit(c1, c2, x1 + up1 - k * normal, normal, dist);
    else emit(c1, c2, x2 + hl2 * d1 + k * normal, normal, dist);
    if (-hl1 > c2x - hl2) emit(c1, c2, x1 - up1 - k * normal, normal, dist);
    else emit(c1, c2, x2 - hl2 * d1 + k * normal, normal, dist);
  } else {
    const double k = r2 + 0.5 * dist;
    emit(c1, c2, cp2 + normal * k, normal, dist);
  }
}

// collideCapsulePlane, MaxContacts.h:2744-2780: each end sphere against the plane; the capsule
// is body1.
template <class Emit>
__device__ void np_capsule_plane(const Bodies& B, uint32_t c, uint32_t p, Emit& emit) {
  const double* R = B.R + 9 * (size_t)c;
  const double rc = B.param[3 * (size_t)c], len = B.param[3 * (size_t)c + 1];
  const d3 c_up = (0.5 * len) * mk3(R[0], R[3], R[6]);
  const d3 cpos = ld3(B.pos, c);
  const d3 posUp = cpos + c_up, posDn = cpos - c_up;
  const d3 n = ld3(B.pnormal, p);
  const double pd = B.pd[p];
  const double distUp = dot3(posUp, n) - pd - rc;
  const double distDn = dot3(posDn, n) - pd - rc;
  if (distUp < kContactThreshold) emit(c, p, posUp - rc * n, n, distUp);
  if (distDn < kContactThreshold) emit(c, p, posDn - rc * n, n, distDn);
}

// A sphere (body or union member) against body b, which is not a union.
template <class Emit>
__device__ void np_sphere_any(const Bodies& B, const SphereRef& s, uint32_t b, Emit& emit) {
  const int tb = B.shape[b];
  if (tb == kShapeSphere) np_sphere_sphere(s, body_sphere(B, b), emit);
  else if (tb == kShapeBox) np_sphere_box(B, s, b, emit);
  else if (tb == kShapeCapsule) np_sphere_capsule(B, s, b, emit);
  else if (tb == kShapePlane) np_sphere_plane(B, s, b, emit);
}

// MaxContacts::collide dispatch, MaxContacts.h:242-526 (sphere/box/capsule/plane and unions of
// spheres).  A union meets another body member by member in the union's order
// (collide{Sphere,Box,Capsule,Plane}Union, MaxContacts.h:843-848, 2576-2581, 2828-2833, 3453-3457);
// two unions pair every member of the one with the smaller SID with every member of the other
// (collideUnionUnion, MaxContacts.h:3569-3581).
template <class Emit>
__device__ void narrowphase(const Bodies& B, uint32_t a, uint32_t b, Emit& emit) {
  const int ta = B.shape[a], tb = B.shape[b];
  if (ta == kShapeUnion || tb == kShapeUnion) {
    if (ta == kShapeUnion && tb == kShapeUnion) {
      if (B.sid[b] < B.sid[a]) { const uint32_t t = a; a = b; b = t; }
      for (uint32_t m1 = B.mstart[a]; m1 < B.mstart[a + 1]; ++m1)
        for (uint32_t m2 = B.mstart[b]; m2 < B.mstart[b + 1]; ++m2)
          np_sphere_sphere(member_sphere(B, m1), member_sphere(B, m2), emit);
    } else {
      const uint32_t u = (ta == kShapeUnion) ? a : b, o = (ta == kShapeUnion) ? b : a;
      for (uint32_t m = B.mstart[u]; m < B.mstart[u + 1]; ++m) np_sphere_any(B, member_sphere(B, m), o, emit);
    }
    return;
  }
  if (ta == kShapeSphere) {
    np_sphere_any(B, body_sphere(B, a), b, emit);
  } else if (ta == kShapeBox) {
    if (tb == kShapeSphere) np_sphere_box(B, body_sphere(B, b), a, emit);
    else if (tb == kShapeBox) np_box_box(B, a, b, emit);
    else if (tb == kShapeCapsule) np_box_capsule(B, a, b, emit);
    else if (tb == kShapePlane) np_box_plane(B, a, b, emit);
  } else if (ta == kShapeCapsule) {
    if (tb == kShapeSphere) np_sphere_capsule(B, body_sphere(B, b), a, emit);
    else if (tb == kShapeBox) np_box_capsule(B, b, a, emit);
    else if (tb == kShapeCapsule) np_capsule_capsule(B, a, b, emit);
    else if (tb == kShapePlane) np_capsule_plane(B, a, b, emit);
  } else {
    if (tb == kShapeSphere) np_sphere_plane(B, body_sphere(B, b), a, emit);
    else if (tb == kShapeBox) np_box_plane(B, b, a, emit);
    else if (tb == kShapeCapsule) np_capsule_plane(B, b, a, emit);
  }
}

// Friction choice of Contact::Contact (src/core/contact/Contact.cpp:83-105): true (dynamic)
// if the tangential relative speed exceeds frictionThreshold, false (static) otherwise.
__device__ inline bool contact_is_dynamic(const Bodies& B, uint32_t b1, uint32_t b2, d3 gpos, d3 n) {
  const d3 x1 = ld3(B.pos, b1), x2 = ld3(B.pos, b2);
  const d3 v1 = ld3(B.v, b1) + cross3(ld3(B.w, b1), gpos - x1);  // RigidBody::velFromWF, RigidBody.h:1080-1085
  const d3 v2 = ld3(B.v, b2) + cross3(ld3(B.w, b2), gpos - x2);
  const d3 rvel = v1 - v2;
  const double nvel = dot3(n, rvel);
  const double tvel = dot3(n, rvel - n * nvel);
  return fabs(tvel) > kFrictionThreshold;
}

// Contact friction from the material pair tables (Materials.h:500-537) of the colliding
// geometries g1, g2; the relative velocity is the top-level bodies' (Contact.cpp:81-105).
__device__ inline double contact_friction(const Bodies& B, const double* csf, const double* cdf,
                                          uint32_t nmat, uint32_t g1, uint32_t g2, d3 gpos, d3 n) {
  const size_t idx = (size_t)geom_mat(B, g1) * nmat + geom_mat(B, g2);
  return contact_is_dynamic(B, geom_body(B, g1), geom_body(B, g2), gpos, n) ? cdf[idx] : csf[idx];
}

// std::max(a, b) as the reference evaluates it: (a < b) ? b : a.
PE_HD double smax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------------------------
// Solver contacts: field-major arrays (field f of contact i at base[f * cap + i]).
// ---------------------------------------------------------------------------------------------
enum ContactField {
  CF_R1 = 0, CF_R2 = 3, CF_N = 6, CF_T = 9, CF_O = 12,
  CF_DDT = 15,   // dist * (1/dt), dist already scaled by erp if negative
  CF_MU = 16, CF_DNI = 17,
  CF_DINV = 18,  // 9: inverse of the 3x3 block in the contact frame
  CF_DTOI = 27,  // 4: inverse of its tangential 2x2 block
  CF_DTO = 31,   // 4: the tangential 2x2 block (golden-section objective)
  CF_P = 35,     // 3: impulse, world frame
  CF_D0 = 38,    // 5: the 3x3 block's first row and column, D00 D01 D02 D10 D20 (projections)
  CF_COUNT = 43
};

struct SolverContacts {
  size_t cap;
  uint32_t* b1;  // body index | kFixedBit
  uint32_t* b2;
  double* f;     // CF_COUNT * cap
};

__device__ __forceinline__ d3 cf_ld3(const double* f, size_t cap, int fld, size_t i) {
  return mk3(f[fld * cap + i], f[(fld + 1) * cap + i], f[(fld + 2) * cap + i]);
}

// Golden-section search of relaxInelasticCoulombContactsByDecoupling (HCSITS.h:2429-2509)
// over the angle of the tangential impulse on the circle of radius flimit.
struct GoldenState {
  double al, ar, am, fm, xm0, xm1;
  bool left;  // "leftlarger"
};

__device__ __forceinline__ double golden_objective(double x0, double x1, double A00, double A01, double A10,
                                                   double A11, double gt, double go) {
  return 0.5 * x0 * (A00 * x0 + A01 * x1) + 0.5 * x1 * (A10 * x0 + A11 * x1) - (x0 * (-gt) + x1 * (-go));
}

// One iteration's update of the bracket given the new point an (with its impulse and objective)
// and the comparison fn < fm.
__device__ __forceinline__ void golden_step(GoldenState& S, double an, double x0, double x1, double fn, bool lt) {
  if (S.left) {
    if (lt) { S.ar = S.am; S.am = an; S.xm0 = x0; S.xm1 = x1; S.fm = fn; S.left = true; }
    else { S.al = an; S.left = false; }
  } else {
    if (lt) { S.al = S.am; S.am = an; S.xm0 = x0; S.xm1 = x1; S.fm = fn; S.left = false; }
    else { S.ar = an; S.left = true; }
  }
}

// Bracket of the tree node `code` (heap index + 1, `depth` levels below the root): the bracket
// updates of golden_step along the node's path, written with selects so that the lanes of a
// wavefront, each on its own path, do not diverge.  LEVELS bounds the depth of the nodes used.
template <int LEVELS>
__device__ __forceinline__ void golden_path(uint32_t code, int depth, double& al, double& ar, double& am,
                                            bool& left) {
#pragma unroll
  for (int lev = 0; lev < LEVELS; ++lev) {
    const bool act = lev < depth;
    const bool lt = ((code >> (act ? depth - 1 - lev : 0)) & 1u) != 0;
    const double a = al + (ar - am);
    // golden_step's four cases, inactive levels leaving everything as it is
    const bool to_am = act && lt;                   // a becomes the interior point
    const bool l_am = to_am && !left, r_am = to_am && left;
    const bool l_a = act && !lt && left, r_a = act && !lt && !left;
    const double nal = l_am ? am : (l_a ? a : al);
    const double nar = r_am ? am : (r_a ? a : ar);
    am = to_am ? a : am;
    al = nal;
    ar = nar;
    left = (act && !lt) ? !left : left;
  }
}

// Value of x in lane `lane` of the wavefront (lane uniform across the wavefront).
__device__ __forceinline__ double wave_read_lane(double x, int lane) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Value of x in lane `node` of the caller's group of 2^SPEC lanes (node uniform in the group).
// UNIFORM: node is the same in the whole wavefront, which reads it with the scalar unit.
template <int SPEC, bool UNIFORM>
__device__ __forceinline__ double group_read(double x, uint32_t node) {
  if constexpr (UNIFORM) return wave_read_lane(x, (int)node);
  else return __shfl(x, (int)node, 1 << SPEC);
}
template <int SPEC, bool UNIFORM>
__device__ __forceinline__ bool group_read(bool x, uint32_t node) {
  if constexpr (UNIFORM) return __builtin_amdgcn_readlane(x ? 1 : 0, (int)node) != 0;
  else return __shfl(x ? 1 : 0, (int)node, 1 << SPEC) != 0;
}

// Points of the decoupled model's search (HCSITS.h:2450-2470): the circle of radius flimit,
// energy 1/2 x^T A x - x^T (-g) of the tangential block.
template <class Trig>
struct CircleEval {
  double flimit, A00, A01, A10, A11, gt, go;
  __device__ __forceinline__ double operator()(double a, double& x0, double& x1) const {
    double sn, cn;
    Trig::sincos(a, sn, cn);
    x0 = flimit * cn;
    x1 = flimit * sn;
    return golden_objective(x0, x1, A00, A01, A10, A11, gt, go);
  }
};

// Points of the generalized-maximum-dissipation search (HCSITS.h:2924-2932): the friction
// cone's boundary under the normal constraint, radius num / (D00 + ma3 cos(a + shiftI)) with
// num = mu (-gdot_n) and ma3 = mu a3 (infinite where negative), energy of the tangential
// block after eliminating the normal row (A_, b_ = -(gt, go)).
template <class Trig>
struct ConeEval {
  double num, D00, ma3, shiftI, A00, A01, A10, A11, gt, go;
  __device__ __forceinline__ double operator()(double a, double& x0, double& x1) const {
    double r_ub = num / (D00 + ma3 * Trig::cos(a + shiftI));
    if (r_ub < 0) r_ub = __builtin_inf();
    double sn, cn;
    Trig::sincos(a, sn, cn);
    x0 = r_ub * cn;
    x1 = r_ub * sn;
    return golden_objective(x0, x1, A00, A01, A10, A11, gt, go);
  }
};

// maxSub iterations of the golden-section search from the bracket in S (S.am its interior
// point); `eval` gives a point's impulse and energy.  SPEC = 0: one lane, as pe (the caller has
// evaluated S.am).  SPEC = d > 0: the 2^d lanes of a group run d iterations per round -- lane
// j evaluates node j of the depth-d tree of possible next points (heap order; the path to a
// node, read from the bits of j + 1 below the leading one, is a sequence of comparison
// outcomes, and the bracket along it depends only on the points, not on the energies).  The
// comparisons are then resolved for all nodes at once: the best value the sequential search
// holds on reaching node j is the running minimum (by the same strict comparison) of the
// energies at j's ancestors, so every lane compares its own value against that fold of its
// ancestors' values, one ballot collects the outcomes, the group walks the tree along them, and
// the bracket after the round is one golden_step from the state at the last node on the walk.
// Every evaluated point and every comparison is the one the sequential search makes, bit for
// bit.  In the first round the group's spare lane evaluates S.am.
//
// SPEC = 6 (a whole wavefront owns the contact) runs the 5-level tree of 32 lanes in both halves
// of the wavefront alike, so the taken node is wave-uniform and read by the scalar unit: rounds of
// a 6-level tree cost enough more (cross-lane reads over 64 lanes) that 20 steps in 4 rounds of 5
// levels take 6.0k cycles against 8.2k in rounds of 6, 6, 6 and 2 (one wavefront, device libm).
template <int SPEC_, class Eval>
__device__ __forceinline__ void golden_section(GoldenState& S, int maxSub, const Eval& eval) {
  constexpr bool UNIFORM = SPEC_ == 6;
  constexpr int SPEC = SPEC_ == 6 ? 5 : SPEC_;
  if constexpr (SPEC == 0) {
    for (int k = 0; k < maxSub; ++k) {
      const double an = S.al + (S.ar - S.am);
      double xn0, xn1;
      const double fn = eval(an, xn0, xn1);
      golden_step(S, an, xn0, xn1, fn, fn < S.fm);
    }
  } else {
    constexpr int L = 1 << SPEC;
    const uint32_t lane = threadIdx.x & (L - 1);
    const uint32_t code = lane + 1;
    const int depth = 31 - __clz(code);   // nodes 0 .. L-2 are used: depth <= SPEC - 1
    const uint32_t base = (threadIdx.x & 63u) & ~(uint32_t)(L - 1);
    for (int k = 0; k < maxSub; k += SPEC) {
      const int levels = maxSub - k < SPEC ? maxSub - k : SPEC;
      double al = S.al, ar = S.ar, am = S.am;
      bool left = S.left;
      golden_path<SPEC - 1>(code, depth, al, ar, am, left);
      // lane L-1 holds no node of the tree: in the first round it evaluates the search's
      // initial point (the bracket's interior point am), which the comparisons need
      const bool init = k == 0 && lane == (uint32_t)(L - 1);
      const double an = init ? S.am : al + (ar - am);
      double xn0, xn1;
      const double fn = eval(an, xn0, xn1);
      if (k == 0) {
        S.fm = group_read<SPEC, UNIFORM>(fn, L - 1);
        S.xm0 = group_read<SPEC, UNIFORM>(xn0, L - 1);
        S.xm1 = group_read<SPEC, UNIFORM>(xn1, L - 1);
      }
      // best value held on reaching this node: fold of the ancestors' values, root first (all
      // the ancestors' values are read before the fold, so the cross-lane reads overlap)
      double fanc[SPEC - 1 > 0 ? SPEC - 1 : 1];
#pragma unroll
      for (int l = 0; l < SPEC - 1; ++l) {
        const uint32_t src = l < depth ? (code >> (depth - l)) - 1u : lane;
        fanc[l] = __shfl(fn, (int)src, L);
      }
      double m = S.fm;
#pragma unroll
      for (int l = 0; l < SPEC - 1; ++l) m = (l < depth && fanc[l] < m) ? fanc[l] : m;
      const unsigned long long all = __ballot(fn < m);
      // (UNIFORM: both halves hold the same tree; the lower half's bits are wave-uniform)
      const uint64_t bits = UNIFORM ? (all & ((1ull << L) - 1ull)) : (all >> base) & ((1ull << L) - 1ull);
      // walk the taken path: fin = the last node compared, last = the last improving one
      uint32_t node = 0, fin = 0;
      int last = -1;
      bool ltf = false;
      for (int d = 0; d < levels; ++d) {
        fin = node;
        ltf = ((bits >> node) & 1ull) != 0;
        if (ltf) last = (int)node;
        node = 2 * node + (ltf ? 2u : 1u);
      }
      const double fal = group_read<SPEC, UNIFORM>(al, fin), far = group_read<SPEC, UNIFORM>(ar, fin);
      const double fam = group_read<SPEC, UNIFORM>(am, fin), fa = group_read<SPEC, UNIFORM>(an, fin);
      const bool fleft = group_read<SPEC, UNIFORM>(left, fin);
      if (fleft) {
        if (ltf) { S.al = fal; S.ar = fam; S.am = fa; S.left = true; }
        else { S.al = fa; S.ar = far; S.am = fam; S.left = false; }
      } else {
        if (ltf) { S.al = fam; S.ar = far; S.am = fa; S.left = false; }
        else { S.al = fal; S.ar = fa; S.am = fam; S.left = true; }
      }
      if (last >= 0) {
        S.fm = group_read<SPEC, UNIFORM>(fn, (uint32_t)last);
        S.xm0 = group_read<SPEC, UNIFORM>(xn0, (uint32_t)last);
        S.xm1 = group_read<SPEC, UNIFORM>(xn1, (uint32_t)last);
      }
    }
  }
}

// One contact's NBGS update: HCSITS.h:2341-2536 (InelasticCoulombContactByDecoupling),
// 2141-2290 (ApproximateInelasticCoulombContactByDecoupling), 2048-2093
// (InelasticFrictionlessContact), 2631-2723 (orthogonal projections, models 2 and 4),
// 2756-3040 (InelasticGeneralizedMaximumDissipationContact, model 5).  Returns |p_old - p_new|_inf (or |p_old|_inf for a
// separating contact).  The caller guarantees that no other thread touches the
// non-fixed bodies of contact i concurrently.  Trig = TrigFast (device libm) or TrigExact
// (correctly rounded, reproduces the host libm's golden-section decisions).
// Loads and stores of the data a sweep writes (velocity corrections, impulses).  COHERENT:
// agent-scope atomic accesses, for sweeps in which consecutive updates of one body may run on
// different compute units (and L2 slices) within one kernel.
template <bool COHERENT>
__device__ __forceinline__ double sweep_ld(const double* p) {
  if constexpr (COHERENT) return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <bool COHERENT>
__device__ __forceinline__ void sweep_st(double* p, double v) {
  if constexpr (COHERENT) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

// Per-body inputs of a relaxation: predicted velocities, inverse mass, world-frame inverse inertia
// (a solver body record, kSB doubles).
struct RelaxBody {
  d3 v, w;
  double im;
  m33 I;
};
__device__ __forceinline__ RelaxBody relax_body(const double* s) {
  RelaxBody b;
  b.v = mk3(s[SB_V], s[SB_V + 1], s[SB_V + 2]);
  b.w = mk3(s[SB_W], s[SB_W + 1], s[SB_W + 2]);
  b.im = s[SB_IM];
#pragma unroll
  for (int k = 0; k < 9; ++k) b.I.a[k] = s[SB_IW + k];
  return b;
}

// Reader of contact i's fields in the field-major solver arrays.
struct GlobalFields {
  const double* f;
  size_t cap, i;
  __device__ __forceinline__ double operator()(int k) const { return f[(size_t)k * cap + i]; }
};

// The arithmetic of one contact's NBGS update, wherever its data live: cf(field) reads the
// contact's fields (CF_*), b1 / b2 are its bodies, dv / dw their velocity corrections (updated in
// place) and p its world-frame impulse (updated in place).  Returns |p_old - p_new|_inf (or
// |p_old|_inf for a separating contact).
template <int MODEL, class Trig, int SPEC, class CF>
__device__ inline double relax_update(const CF& cf, const RelaxBody& b1, const RelaxBody& b2, d3& dv1, d3& dw1,
                                      d3& dv2, d3& dw2, d3& p, int maxSub) {
  const d3 v1 = b1.v, w1 = b1.w, v2 = b2.v, w2 = b2.w;
  const double im1 = b1.im, im2 = b2.im;
  const m33& I1 = b1.I;
  const m33& I2 = b2.I;
  const d3 r1 = mk3(cf(CF_R1), cf(CF_R1 + 1), cf(CF_R1 + 2)), r2 = mk3(cf(CF_R2), cf(CF_R2 + 1), cf(CF_R2 + 2));
  const d3 n = mk3(cf(CF_N), cf(CF_N + 1), cf(CF_N + 2)), t = mk3(cf(CF_T), cf(CF_T + 1), cf(CF_T + 2));
  const d3 o = mk3(cf(CF_O), cf(CF_O + 1), cf(CF_O + 2));
  const double ddt = cf(CF_DDT);

  // Remove this contact's reaction (HCSITS.h:2343-2346).
  dv1 = dv1 - im1 * p;
  dw1 = dw1 - mat_vec(I1, cross3(r1, p));
  dv2 = dv2 + im2 * p;
  dw2 = dw2 + mat_vec(I2, cross3(r2, p));

  // Relative contact velocity in the contact frame (HCSITS.h:2349-2360).
  const d3 gdot = (v1 + dv1) - (v2 + dv2) + cross3(w1 + dw1, r1) - cross3(w2 + dw2, r2);
  const m33 cfm = mat_cols(n, t, o);
  d3 g = mat_tvec(cfm, gdot);
  g.x += ddt;

  double delta;
  if constexpr (MODEL == 2 || MODEL == 4) {
    // Orthogonal projection onto the friction cone (HCSITS.h:2631-2723): one step from the
    // current impulse, p_cf - (D p_cf + gdot_nto), Anitescu/Tasora (2); De Saxce/Feng (4) adds
    // mu |(D p_cf + gdot_nto)_to| to the normal component first.  pe's branches are kept as
    // written: its separating test compares fsq with p_n / mu^2 < 0 and so never holds.
    const double mu = cf(CF_MU);
    const double D00 = cf(CF_D0), D01 = cf(CF_D0 + 1), D02 = cf(CF_D0 + 2);
    const double D10 = cf(CF_D0 + 3), D20 = cf(CF_D0 + 4);
    const double D11 = cf(CF_DTO), D12 = cf(CF_DTO + 1);
    const double D21 = cf(CF_DTO + 2), D22 = cf(CF_DTO + 3);
    const d3 pc = mat_tvec(cfm, p);
    d3 tmp = mk3((D00 * pc.x + D01 * pc.y + D02 * pc.z) + g.x, (D10 * pc.x + D11 * pc.y + D12 * pc.z) + g.y,
                 (D20 * pc.x + D21 * pc.y + D22 * pc.z) + g.z);
    if (MODEL == 4) tmp.x += sqrt(tmp.y * tmp.y + tmp.z * tmp.z) * mu;
    d3 pcf = pc - tmp;
    const double flimit = mu * pcf.x;
    const double fsq = pcf.y * pcf.y + pcf.z * pcf.z;
    if (pcf.x > 0 && fsq < flimit * flimit) {
      // inside the cone: static
    } else if (pcf.x < 0 && fsq < pcf.x / (mu * mu)) {
      pcf = mk3(0, 0, 0);
    } else {
      const double fn = sqrt(fsq);
      pcf.x = (fn * mu + pcf.x) / (mu * mu + 1.0);
      const double factor = mu * pcf.x / fn;
      pcf.y *= factor;
      pcf.z *= factor;
    }
    const d3 pwf = mat_vec(cfm, pcf);
    const d3 dp = p - pwf;
    delta = fmax(fabs(dp.x), fmax(fabs(dp.y), fabs(dp.z)));
    p = pwf;
    dv1 = dv1 + im1 * p;
    dw1 = dw1 + mat_vec(I1, cross3(r1, p));
    dv2 = dv2 - im2 * p;
    dw2 = dw2 - mat_vec(I2, cross3(r2, p));
  } else if (g.x >= 0) {
    delta = fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z)));
    p = mk3(0, 0, 0);
  } else {
    d3 pwf;
    if (MODEL == 0) {
      pwf = n * (-cf(CF_DNI) * g.x);
    } else {
      m33 Dinv;
#pragma unroll
      for (int k = 0; k < 9; ++k) Dinv.a[k] = cf(CF_DINV + k);
      const double mu = cf(CF_MU);
      double pc[3];
      {
        const d3 s = -mat_vec(Dinv, g);
        pc[0] = s.x; pc[1] = s.y; pc[2] = s.z;
      }
      double flimit = mu * pc[0];
      double fsq = pc[1] * pc[1] + pc[2] * pc[2];
      if (fsq > flimit * flimit || pc[0] < 0) {
        const double dni = cf(CF_DNI);
        const double T00 = cf(CF_DTOI + 0), T01 = cf(CF_DTOI + 1);
        const double T10 = cf(CF_DTOI + 2), T11 = cf(CF_DTOI + 3);
        if (MODEL == 1) {
          // ApproximateInelasticCoulombContactByDecoupling, HCSITS.h:2186-2226
          d3 ptmp = dot3(t, p) * t + dot3(o, p) * o;
          const double gn = g.x + dot3(n, cross3(mat_vec(I1, cross3(r1, ptmp)), r1) +
                                             cross3(mat_vec(I2, cross3(r2, ptmp)), r2));
          pc[0] = -(dni * gn);
          pc[0] = smax(0.0, pc[0]);
          ptmp = n * pc[0];
          const d3 g2 = gdot + cross3(mat_vec(I1, cross3(r1, ptmp)), r1) + cross3(mat_vec(I2, cross3(r2, ptmp)), r2);
          const double gt = dot3(t, g2), go = dot3(o, g2);
          pc[1] = -(T00 * gt + T01 * go);
          pc[2] = -(T10 * gt + T11 * go);
          flimit = mu * pc[0];
          fsq = pc[1] * pc[1] + pc[2] * pc[2];
          if (fsq > flimit * flimit) {
            const double fr = flimit / sqrt(fsq);
            pc[1] *= fr;
            pc[2] *= fr;
          }
        } else if constexpr (MODEL == 5) {
          // InelasticGeneralizedMaximumDissipationContact, HCSITS.h:2806-3022: the impulse on
          // the friction cone's boundary that meets the normal constraint with the least energy
          // of its tangential part, by golden section over the angle, bracketed where the
          // boundary's radius is positive (angleI) and cheaper than the unconstrained direction
          // (angleJ); the unconstrained direction replaces the result if cheaper.
          constexpr double kTwoPi = 2.0 * 3.141592653589793;   // 2 * M_PI, pe/math/Constants.h:163
          const double D00 = cf(CF_D0), D01 = cf(CF_D0 + 1), D02 = cf(CF_D0 + 2);
          const double D11 = cf(CF_DTO), D12 = cf(CF_DTO + 1);
          const double D22 = cf(CF_DTO + 3);
          const double invA_nn = 1.0 / D00;
          const double offdiag = D12 - invA_nn * D01 * D02;
          const double A00 = D11 - invA_nn * (D01 * D01), A11 = D22 - invA_nn * (D02 * D02);
          const double bb0 = -g.y + invA_nn * D01 * g.x;
          const double bb1 = -g.z + invA_nn * D02 * g.x;
          const double shiftI = Trig::atan2(-D02, D01);
          const double shiftJ = Trig::atan2(-pc[2], pc[1]);
          const double a3 = sqrt(D01 * D01 + D02 * D02);
          const double fractionI = -D00 / (mu * a3);
          const double fj = invA_nn * mu * ((-g.x) / sqrt(fsq) - a3 * Trig::cos(shiftI - shiftJ));
          const double fractionJ = (1.0 < fj) ? 1.0 : fj;   // std::min(fj, 1)
          GoldenState gs;
          if (fractionJ < -1) {
            const double angleI = Trig::acos(fractionI);
            gs.al = -angleI - shiftI;
            gs.ar = +angleI - shiftI;
            if (gs.al < 0) { gs.al += kTwoPi; gs.ar += kTwoPi; }
          } else if (D00 > mu * a3) {
            const double angleJ = Trig::acos(fractionJ);
            gs.al = -angleJ - shiftJ;
            gs.ar = +angleJ - shiftJ;
            if (gs.al < 0) { gs.al += kTwoPi; gs.ar += kTwoPi; }
          } else {
            const double angleJ = Trig::acos(fractionJ);
            double a1l = -angleJ - shiftJ, a1r = +angleJ - shiftJ;
            if (a1l < 0) { a1l += kTwoPi; a1r += kTwoPi; }
            const double angleI = Trig::acos(fractionI);
            double a2l = -angleI - shiftI, a2r = +angleI - shiftI;
            if (a2l < 0) { a2l += kTwoPi; a2r += kTwoPi; }
            if (a1l > a2l) {
              const double tl = a1l, tr = a1r;
              a1l = a2l; a1r = a2r; a2l = tl; a2r = tr;
            }
            if (a2l > a1r) {
              a2r -= kTwoPi;
              if (!(a2r > a1r)) a1r = a2r;
            } else {
              a1l = a2l;
              if (!(a2r > a1r)) a1r = a2r;
            }
            gs.al = a1l;
            gs.ar = a1r;
          }
          const double phi = 0.5 * (1.0 + sqrt(5.0));
          gs.am = (gs.ar + gs.al * phi) / (1 + phi);
          gs.left = false;
          const ConeEval<Trig> ev{mu * (-g.x), D00, mu * a3, shiftI, A00, offdiag, offdiag, A11, -bb0, -bb1};
          // (with SPEC > 0 the search's first round evaluates the initial point)
          if (SPEC == 0 || maxSub <= 0) gs.fm = ev(gs.am, gs.xm0, gs.xm1);
          golden_section<SPEC>(gs, maxSub, ev);
          {
            double xi0, xi1;
            const double f_init = ev(Trig::atan2(pc[2], pc[1]), xi0, xi1);
            if (f_init < gs.fm) { gs.xm0 = xi0; gs.xm1 = xi1; }
          }
          pc[0] = invA_nn * (-g.x - D01 * gs.xm0 - D02 * gs.xm1);
          pc[1] = gs.xm0;
          pc[2] = gs.xm1;
        } else {
          // InelasticCoulombContactByDecoupling: 20 decoupled passes, HCSITS.h:2395-2520.
          // Pass j computes the normal impulse c_j from the previous pass's tangential impulse
          // t_{j-1} and then t_j from c_j alone, so the passes are a deterministic map of c.
          // Once c_j equals c_{j-P} bit for bit (P = 1, 2, 3) the passes repeat with period P
          // from pass j-P on, the result of pass 19 is that of pass j-P+((19-j+P) mod P), and
          // the remaining passes are skipped: the same bits as running all 20.
          const double A00 = cf(CF_DTO + 0), A01 = cf(CF_DTO + 1);
          const double A10 = cf(CF_DTO + 2), A11 = cf(CF_DTO + 3);
          double c1 = 0.0, c2 = 0.0, c3 = 0.0;  // c_{j-1}, c_{j-2}, c_{j-3}
          double u1 = 0.0, u2 = 0.0;            // t_{j-2}
          double w1 = 0.0, w2 = 0.0;            // t_{j-3}
          for (int j = 0; j < 20; ++j) {
            const double tq1 = pc[1], tq2 = pc[2];  // t_{j-1}
            pc[0] = 0;
            const d3 tp = t * pc[1] + o * pc[2];
            d3 gc = gdot + cross3(mat_vec(I1, cross3(r1, tp)), r1) + cross3(mat_vec(I2, cross3(r2, tp)), r2);
            const double gcn = dot3(n, gc) + ddt;
            pc[0] = smax(0.0, -(dni * gcn));
            const long long cb = __double_as_longlong(pc[0]);
            if (j >= 1 && cb == __double_as_longlong(c1)) break;  // pass 19 = pass j: (c_j, t_{j-1})
            if (j >= 2 && cb == __double_as_longlong(c2)) {
              if (((19 - j) & 1) == 0) { pc[1] = u1; pc[2] = u2; }  // pass j: (c_j, t_{j-2})
              else { pc[0] = c1; }                                  // pass j-1: (c_{j-1}, t_{j-1})
              break;
            }
            if (j >= 3 && cb == __double_as_longlong(c3)) {
              const int r = (19 - j) % 3;
              if (r == 0) { pc[1] = w1; pc[2] = w2; }                   // pass j: (c_j, t_{j-3})
              else if (r == 1) { pc[0] = c2; pc[1] = u1; pc[2] = u2; }  // pass j-2: (c_{j-2}, t_{j-2})
              else { pc[0] = c1; }                                      // pass j-1: (c_{j-1}, t_{j-1})
              break;
            }
            c3 = c2; c2 = c1; c1 = pc[0];
            w1 = u1; w2 = u2;
            u1 = tq1; u2 = tq2;
            pc[1] = pc[2] = 0.0;
            const d3 np0 = n * pc[0];
            gc = gdot + cross3(mat_vec(I1, cross3(r1, np0)), r1) + cross3(mat_vec(I2, cross3(r2, np0)), r2);
            const double gt = dot3(t, gc), go = dot3(o, gc);
            pc[1] = -(T00 * gt + T01 * go);
            pc[2] = -(T10 * gt + T11 * go);
            flimit = mu * pc[0];
            fsq = pc[1] * pc[1] + pc[2] * pc[2];
            if (fsq > flimit * flimit) {
              const double fr = flimit / sqrt(fsq);
              const double phi = 0.5 * (1.0 + sqrt(5.0));
              const double shift = Trig::atan2(-pc[2], pc[1]);
              const double acos_f = Trig::acos(fr);
              GoldenState gs;
              gs.al = -acos_f - shift;
              gs.ar = acos_f - shift;
              gs.am = (gs.ar + gs.al * phi) / (1 + phi);
              gs.left = false;
              const CircleEval<Trig> ev{flimit, A00, A01, A10, A11, gt, go};
              // (with SPEC > 0 the search's first round evaluates the initial point)
              if (SPEC == 0 || maxSub <= 0) gs.fm = ev(gs.am, gs.xm0, gs.xm1);
              golden_section<SPEC>(gs, maxSub, ev);
              pc[1] = gs.xm0;
              pc[2] = gs.xm1;
            }
          }
        }
      }
      pwf = mat_vec(cfm, mk3(pc[0], pc[1], pc[2]));
    }
    const d3 dp = p - pwf;
    delta = fmax(fabs(dp.x), fmax(fabs(dp.y), fabs(dp.z)));
    p = pwf;
    // Apply the new reaction (HCSITS.h:2532-2535).
    dv1 = dv1 + im1 * p;
    dw1 = dw1 + mat_vec(I1, cross3(r1, p));
    dv2 = dv2 - im2 * p;
    dw2 = dw2 - mat_vec(I2, cross3(r2, p));
  }
  return delta;
}

// e1, e2: the contact's body entries C.b1[i], C.b2[i] (the caller has them already).
// COHERENT: the whole wavefront owns contact i (the dataflow kernel).
template <int MODEL, class Trig, int SPEC = 0, bool COHERENT = false>
__device__ inline double relax_contact(const SolverContacts& C, size_t i, uint32_t e1, uint32_t e2,
                                       const double* __restrict__ sb, double* __restrict__ dvw, double dtinv,
                                       int maxSub);
template <int MODEL, class Trig, int SPEC = 0, bool COHERENT = false>
__device__ inline double relax_contact(const SolverContacts& C, size_t i, const double* __restrict__ sb,
                                       double* __restrict__ dvw, double dtinv, int maxSub) {
  return relax_contact<MODEL, Trig, SPEC, COHERENT>(C, i, C.b1[i], C.b2[i], sb, dvw, dtinv, maxSub);
}
template <int MODEL, class Trig, int SPEC, bool COHERENT>
__device__ inline double relax_contact(const SolverContacts& C, size_t i, uint32_t e1, uint32_t e2,
                                       const double* __restrict__ sb, double* __restrict__ dvw, double dtinv,
                                       int maxSub) {
  (void)dtinv;  // the contact record holds dist / dt
  const size_t cap = C.cap;
  double* f = C.f;
  const uint32_t i1 = e1 & ~kFixedBit, i2 = e2 & ~kFixedBit;
  const bool fx1 = (e1 & kFixedBit) != 0, fx2 = (e2 & kFixedBit) != 0;
  const RelaxBody B1 = relax_body(sb + (size_t)kSB * i1);
  const RelaxBody B2 = relax_body(sb + (size_t)kSB * i2);
  double* q1 = dvw + (size_t)kDVW * i1;
  double* q2 = dvw + (size_t)kDVW * i2;
  d3 dv1 = mk3(0, 0, 0), dw1 = mk3(0, 0, 0), dv2 = mk3(0, 0, 0), dw2 = mk3(0, 0, 0);
  if (!fx1) {
    dv1 = mk3(sweep_ld<COHERENT>(q1), sweep_ld<COHERENT>(q1 + 1), sweep_ld<COHERENT>(q1 + 2));
    dw1 = mk3(sweep_ld<COHERENT>(q1 + 3), sweep_ld<COHERENT>(q1 + 4), sweep_ld<COHERENT>(q1 + 5));
  }
  if (!fx2) {
    dv2 = mk3(sweep_ld<COHERENT>(q2), sweep_ld<COHERENT>(q2 + 1), sweep_ld<COHERENT>(q2 + 2));
    dw2 = mk3(sweep_ld<COHERENT>(q2 + 3), sweep_ld<COHERENT>(q2 + 4), sweep_ld<COHERENT>(q2 + 5));
  }
  d3 p;
  if constexpr (COHERENT) {
    // one load instruction for the three components (lanes 0-2), then broadcast: issued one
    // after another, the three agent-scope loads cost three memory round trips
    const uint32_t lane = threadIdx.x & 63u;
    double pl = 0.0;
    if (lane < 3) pl = sweep_ld<true>(f + (CF_P + lane) * cap + i);
    p = mk3(wave_read_lane(pl, 0), wave_read_lane(pl, 1), wave_read_lane(pl, 2));
  } else {
    p = mk3(f[CF_P * cap + i], f[(CF_P + 1) * cap + i], f[(CF_P + 2) * cap + i]);
  }
  const GlobalFields cf{f, cap, i};
  const double delta = relax_update<MODEL, Trig, SPEC>(cf, B1, B2, dv1, dw1, dv2, dw2, p, maxSub);
  // one lane of the contact's group stores (COHERENT: the whole wavefront owns the contact)
  const uint32_t group_mask = COHERENT ? 63u : ((1u << SPEC) - 1u);
  if ((threadIdx.x & group_mask) == 0) {
    sweep_st<COHERENT>(f + CF_P * cap + i, p.x);
    sweep_st<COHERENT>(f + (CF_P + 1) * cap + i, p.y);
    sweep_st<COHERENT>(f + (CF_P + 2) * cap + i, p.z);
    if (!fx1) {
      sweep_st<COHERENT>(q1, dv1.x); sweep_st<COHERENT>(q1 + 1, dv1.y); sweep_st<COHERENT>(q1 + 2, dv1.z);
      sweep_st<COHERENT>(q1 + 3, dw1.x); sweep_st<COHERENT>(q1 + 4, dw1.y); sweep_st<COHERENT>(q1 + 5, dw1.z);
    }
    if (!fx2) {
      sweep_st<COHERENT>(q2, dv2.x); sweep_st<COHERENT>(q2 + 1, dv2.y); sweep_st<COHERENT>(q2 + 2, dv2.z);
      sweep_st<COHERENT>(q2 + 3, dw2.x); sweep_st<COHERENT>(q2 + 4, dw2.y); sweep_st<COHERENT>(q2 + 5, dw2.z);
    }
  }
  return delta;
}

}  // namespace pe_amd
