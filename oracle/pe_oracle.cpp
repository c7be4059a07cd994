/*
 * CPU oracle for the HCSITS rigid-body time step -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded restatement of the pe reference's time step, used only by
 * the tests, the smoke check and the CPU-baseline leg of the benchmark.  The
 * floating-point operation order of the reference is preserved everywhere
 * (it is compiled with -ffp-contract=off and without fast-math), because the
 * HIP path is checked against it to 1e-12.  Citations are file:line in the pe
 * source tree.
 */
#include "oracle_api.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <list>
#include <stdexcept>
#include <vector>

#include <quadmath.h>

// ORC_PLANT selects a deliberate mistake, compiled only by the tests that show their checks catch
// it (tests/test_oracle.py builds such variants into a temporary directory): 0 = none (always the
// case for oracle/_build), 1 = gyroscopic term with the cross product's operands swapped,
// 2 = world-frame inertia as R^T I R instead of R I R^T, 3 = gyroscopic term left out.
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

namespace {

// ---------------------------------------------------------------------------------------------
// Constants.  pe/core/Thresholds.h:147-170 (double), pe/util/Limits.h:385 (accuracy),
// pe/config/HashGrids.h (grid constants).
// ---------------------------------------------------------------------------------------------
const double kContactThreshold = 1e-8;
const double kFrictionThreshold = 1e-8;
const double kAccuracy = 1e-8;
const double kParallelThreshold = 1e-8;   // pe/core/Thresholds.h:170
const size_t kCellCount = 16;
const size_t kMinimalGridDensity = 8;
const size_t kGridActivationThreshold = 32;
const double kHierarchyFactor = 2.0;
const double kInf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------------------------
// Math primitives in the reference's operation order (pe/math/*.h).
// ---------------------------------------------------------------------------------------------
struct V3 {
  double x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double& operator[](int i) { return (&x)[i]; }
  double operator[](int i) const { return (&x)[i]; }
};
inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(const V3& a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(const V3& a, double s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, const V3& a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3& operator+=(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, const V3& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
// trans(a)*b, Vector3.h:2096-2099
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// a % b, Vector3.h:2339-2345
inline V3 cross(const V3& a, const V3& b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Vector3::length, Vector3.h:952-959
inline double length(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3::normalize / getNormalized, Vector3.h:994-1039 (1/len, then multiply)
inline V3 normalized(const V3& a) {
  const double len = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  if (len == 0.0) return a;
  const double ilen = 1.0 / len;
  return V3(a.x * ilen, a.y * ilen, a.z * ilen);
}

struct M3 {
  double m[9];
  M3() { for (int i = 0; i < 9; ++i) m[i] = 0; }
  M3(double a, double b, double c, double d, double e, double f, double g, double h, double i) {
    m[0] = a; m[1] = b; m[2] = c; m[3] = d; m[4] = e; m[5] = f; m[6] = g; m[7] = h; m[8] = i;
  }
  // Column constructor, Matrix3x3.h:351-358
  static M3 cols(const V3& c1, const V3& c2, const V3& c3) {
    return M3(c1[0], c2[0], c3[0], c1[1], c2[1], c3[1], c1[2], c2[2], c3[2]);
  }
  double& operator[](int i) { return m[i]; }
  double operator[](int i) const { return m[i]; }
};
// Matrix3x3.h:2241-2253
inline M3 operator*(const M3& l, const M3& r) {
  return M3(l[0] * r[0] + l[1] * r[3] + l[2] * r[6], l[0] * r[1] + l[1] * r[4] + l[2] * r[7],
            l[0] * r[2] + l[1] * r[5] + l[2] * r[8], l[3] * r[0] + l[4] * r[3] + l[5] * r[6],
            l[3] * r[1] + l[4] * r[4] + l[5] * r[7], l[3] * r[2] + l[4] * r[5] + l[5] * r[8],
            l[6] * r[0] + l[7] * r[3] + l[8] * r[6], l[6] * r[1] + l[7] * r[4] + l[8] * r[7],
            l[6] * r[2] + l[7] * r[5] + l[8] * r[8]);
}
// Matrix3x3.h:2172-2178
inline V3 operator*(const M3& l, const V3& r) {
  return V3(l[0] * r[0] + l[1] * r[1] + l[2] * r[2], l[3] * r[0] + l[4] * r[1] + l[5] * r[2],
            l[6] * r[0] + l[7] * r[1] + l[8] * r[2]);
}
inline M3 operator+(const M3& a, const M3& b) {
  M3 r; for (int i = 0; i < 9; ++i) r[i] = a[i] + b[i]; return r;
}
inline M3 operator-(const M3& a) {
  M3 r; for (int i = 0; i < 9; ++i) r[i] = -a[i]; return r;
}
// trans(), Matrix3x3.h:1857-1860
inline M3 trans(const M3& a) { return M3(a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]); }
// vec % mat = [vec]x * mat, Matrix3x3.h:2417-2430
inline M3 skew_mul(const V3& v, const M3& m) {
  return M3(v[1] * m[6] - v[2] * m[3], v[1] * m[7] - v[2] * m[4], v[1] * m[8] - v[2] * m[5],
            v[2] * m[0] - v[0] * m[6], v[2] * m[1] - v[0] * m[7], v[2] * m[2] - v[0] * m[8],
            v[0] * m[3] - v[1] * m[0], v[0] * m[4] - v[1] * m[1], v[0] * m[5] - v[1] * m[2]);
}
// mat % vec = mat * [vec]x, Matrix3x3.h:2355-2367
inline M3 mul_skew(const M3& m, const V3& v) {
  return M3(m[1] * v[2] - m[2] * v[1], m[2] * v[0] - m[0] * v[2], m[0] * v[1] - m[1] * v[0],
            m[4] * v[2] - m[5] * v[1], m[5] * v[0] - m[3] * v[2], m[3] * v[1] - m[4] * v[0],
            m[7] * v[2] - m[8] * v[1], m[8] * v[0] - m[6] * v[2], m[6] * v[1] - m[7] * v[0]);
}
// Matrix3x3::getInverse, Matrix3x3.h:1084-1104 (cofactors)
inline M3 inverse(const M3& v) {
  double det = v[0] * ((v[4] * v[8]) - (v[7] * v[5])) + v[1] * ((v[5] * v[6]) - (v[8] * v[3])) +
               v[2] * ((v[3] * v[7]) - (v[4] * v[6]));
  det = 1.0 / det;
  return M3(det * ((v[4] * v[8]) - (v[5] * v[7])), det * ((v[7] * v[2]) - (v[8] * v[1])),
            det * ((v[1] * v[5]) - (v[2] * v[4])), det * ((v[5] * v[6]) - (v[3] * v[8])),
            det * ((v[8] * v[0]) - (v[6] * v[2])), det * ((v[2] * v[3]) - (v[0] * v[5])),
            det * ((v[3] * v[7]) - (v[4] * v[6])), det * ((v[6] * v[1]) - (v[7] * v[0])),
            det * ((v[0] * v[4]) - (v[1] * v[3])));
}

struct M2 { double m[4]; };
// Matrix2x2::getInverse, Matrix2x2.h:977-988
inline M2 inverse2(const M2& v) {
  double det = v.m[0] * v.m[3] - v.m[1] * v.m[2];
  det = 1.0 / det;
  M2 r; r.m[0] = det * v.m[3]; r.m[1] = -det * v.m[1]; r.m[2] = -det * v.m[2]; r.m[3] = det * v.m[0];
  return r;
}

struct Q { double v[4]; };
inline Q quat_identity() { Q q; q.v[0] = 1; q.v[1] = 0; q.v[2] = 0; q.v[3] = 0; return q; }
// Quaternion(axis, angle), Quaternion.h:274-294
// Trig of the step: the C library's (as pe), or correctly rounded (evaluated in 113-bit quad
// precision and rounded once to double) when cr is set -- an ideal libm that the device's
// correctly rounded trig is checked against.
inline double t_sin(double x, bool cr) { return cr ? (double)sinq((__float128)x) : std::sin(x); }
inline double t_cos(double x, bool cr) { return cr ? (double)cosq((__float128)x) : std::cos(x); }
inline double t_atan2(double y, double x, bool cr) {
  return cr ? (double)atan2q((__float128)y, (__float128)x) : std::atan2(y, x);
}
inline double t_acos(double x, bool cr) { return cr ? (double)acosq((__float128)x) : std::acos(x); }

inline Q quat_axis_angle(V3 axis, double angle, bool cr = false) {
  if (std::fabs(angle) < 1e-15) return quat_identity();
  const double sina = t_sin(angle * 0.5, cr);
  const double cosa = t_cos(angle * 0.5, cr);
  axis = normalized(axis);
  Q q; q.v[0] = cosa; q.v[1] = sina * axis[0]; q.v[2] = sina * axis[1]; q.v[3] = sina * axis[2];
  return q;
}
// operator*(Quaternion, Quaternion) with conditional renormalisation, Quaternion.h:1146-1160
inline Q quat_mul(const Q& l, const Q& r) {
  const double a = l.v[0] * r.v[0] - l.v[1] * r.v[1] - l.v[2] * r.v[2] - l.v[3] * r.v[3];
  const double i = l.v[0] * r.v[1] + l.v[1] * r.v[0] + l.v[2] * r.v[3] - l.v[3] * r.v[2];
  const double j = l.v[0] * r.v[2] + l.v[2] * r.v[0] + l.v[3] * r.v[1] - l.v[1] * r.v[3];
  const double k = l.v[0] * r.v[3] + l.v[3] * r.v[0] + l.v[1] * r.v[2] - l.v[2] * r.v[1];
  const double len2 = a * a + i * i + j * j + k * k;
  Q q;
  if (std::fabs(len2 - 1.0) < 1e-8) {
    q.v[0] = a; q.v[1] = i; q.v[2] = j; q.v[3] = k;
  } else {
    const double ilen = 1.0 / std::sqrt(len2);
    q.v[0] = a * ilen; q.v[1] = i * ilen; q.v[2] = j * ilen; q.v[3] = k * ilen;
  }
  return q;
}
// Quaternion::toRotationMatrix, Quaternion.h:591-602
inline M3 quat_to_R(const Q& q) {
  const double* v = q.v;
  return M3(1.0 - 2.0 * v[2] * v[2] - 2.0 * v[3] * v[3], 2.0 * (v[1] * v[2] - v[0] * v[3]),
            2.0 * (v[1] * v[3] + v[0] * v[2]), 2.0 * (v[1] * v[2] + v[0] * v[3]),
            1.0 - 2.0 * v[1] * v[1] - 2.0 * v[3] * v[3], 2.0 * (v[2] * v[3] - v[0] * v[1]),
            2.0 * (v[1] * v[3] - v[0] * v[2]), 2.0 * (v[2] * v[3] + v[0] * v[1]),
            1.0 - 2.0 * v[1] * v[1] - 2.0 * v[2] * v[2]);
}
// Quaternion::rotate, Quaternion.h:744-756
inline V3 quat_rotate(const Q& q, const V3& x) {
  const double* v = q.v;
  const double w = v[1] * x[0] + v[2] * x[1] + v[3] * x[2];
  const double a = v[0] * x[0] - v[3] * x[1] + v[2] * x[2];
  const double b = v[0] * x[1] - v[1] * x[2] + v[3] * x[0];
  const double c = v[0] * x[2] - v[2] * x[0] + v[1] * x[1];
  return V3(v[0] * a + v[1] * w + v[2] * c - v[3] * b, v[0] * b + v[2] * w + v[3] * a - v[1] * c,
            v[0] * c + v[3] * w + v[1] * b - v[2] * a);
}

// ---------------------------------------------------------------------------------------------
// Materials: density and the additive friction tables, Materials.cpp:175-247.
// ---------------------------------------------------------------------------------------------
struct Material { double density, cor, csf, cdf; };

// ---------------------------------------------------------------------------------------------
// Rigid bodies (RigidBodyBase.h:131-165; SphereBase, BoxBase, PlaneBase).
// ---------------------------------------------------------------------------------------------
struct Body {
  int shape = 0;
  uint64_t sid = 0;
  int mat = 0;
  bool fixed = false;
  bool awake = true;
  V3 gpos, v, w, force, torque;
  Q q = quat_identity();
  M3 R = M3(1, 0, 0, 0, 1, 0, 0, 0, 1);
  double mass = 0, invMass = 0;
  M3 I, Iinv;                 // body frame
  double radius = 0;          // sphere, capsule
  double length = 0;          // capsule: length of the cylinder part, along the body x-axis
  V3 lengths;                 // box
  V3 normal; double d = 0;    // plane
  int super = -1;             // union member: index of its union in the body storage
  V3 rpos;                    // union member: position in the union's body frame
  int mbegin = 0, mcount = 0; // union: its members, orc_world::members[mbegin, mbegin + mcount)
  double aabb[6] = {0, 0, 0, 0, 0, 0};
  // HashGrids bookkeeping (BodyTrait<HashGrids>): grid, hash, cell id
  int grid = -1;
  size_t hash = 0, cellId = 0;
  bool finite() const { return shape != ORC_PLANE; }

  // RigidBody::getInvInertia / getInertia, RigidBody.h:842-869
#if ORC_PLANT == 2
  M3 invInertiaWF() const { return trans(R) * Iinv * R; }
  M3 inertiaWF() const { return trans(R) * I * R; }
#else
  M3 invInertiaWF() const { return R * Iinv * trans(R); }
  M3 inertiaWF() const { return R * I * trans(R); }
#endif
  // RigidBody::velFromWF, RigidBody.h:1080-1085
  V3 velFromWF(const V3& p) const { return v + cross(w, p - gpos); }

  void calcBoundingBox() {
    if (shape == ORC_SPHERE) {  // SphereBase.h:208-221
      const double len = radius + kContactThreshold;
      aabb[0] = gpos[0] - len; aabb[1] = gpos[1] - len; aabb[2] = gpos[2] - len;
      aabb[3] = gpos[0] + len; aabb[4] = gpos[1] + len; aabb[5] = gpos[2] + len;
    } else if (shape == ORC_BOX) {  // BoxBase.cpp:127-141
      const double xl = 0.5 * (std::fabs(R[0] * lengths[0]) + std::fabs(R[1] * lengths[1]) + std::fabs(R[2] * lengths[2])) + kContactThreshold;
      const double yl = 0.5 * (std::fabs(R[3] * lengths[0]) + std::fabs(R[4] * lengths[1]) + std::fabs(R[5] * lengths[2])) + kContactThreshold;
      const double zl = 0.5 * (std::fabs(R[6] * lengths[0]) + std::fabs(R[7] * lengths[1]) + std::fabs(R[8] * lengths[2])) + kContactThreshold;
      aabb[0] = gpos[0] - xl; aabb[1] = gpos[1] - yl; aabb[2] = gpos[2] - zl;
      aabb[3] = gpos[0] + xl; aabb[4] = gpos[1] + yl; aabb[5] = gpos[2] + zl;
    } else if (shape == ORC_CAPSULE) {  // CapsuleBase.cpp:124-138
      const double xl = std::fabs(R[0] * length) * 0.5 + radius + kContactThreshold;
      const double yl = std::fabs(R[3] * length) * 0.5 + radius + kContactThreshold;
      const double zl = std::fabs(R[6] * length) * 0.5 + radius + kContactThreshold;
      aabb[0] = gpos[0] - xl; aabb[1] = gpos[1] - yl; aabb[2] = gpos[2] - zl;
      aabb[3] = gpos[0] + xl; aabb[4] = gpos[1] + yl; aabb[5] = gpos[2] + zl;
    } else {  // PlaneBase.cpp:132-148
      const V3& n = normal;
      aabb[0] = (n[0] < 0 && n[1] == 0 && n[2] == 0) ? (-d - kContactThreshold) : -kInf;
      aabb[1] = (n[0] == 0 && n[1] < 0 && n[2] == 0) ? (-d - kContactThreshold) : -kInf;
      aabb[2] = (n[0] == 0 && n[1] == 0 && n[2] < 0) ? (-d - kContactThreshold) : -kInf;
      aabb[3] = (n[0] > 0 && n[1] == 0 && n[2] == 0) ? (d + kContactThreshold) : kInf;
      aabb[4] = (n[0] == 0 && n[1] > 0 && n[2] == 0) ? (d + kContactThreshold) : kInf;
      aabb[5] = (n[0] == 0 && n[1] == 0 && n[2] > 0) ? (d + kContactThreshold) : kInf;
    }
  }
  // BodyTrait<HashGrids>::getAABBSize, coarse/BodyTrait.h:209-218
  double aabbSize() const {
    double size = aabb[3] - aabb[0];
    const double yd = aabb[4] - aabb[1];
    const double zd = aabb[5] - aabb[2];
    if (yd > size) size = yd;
    if (zd > size) size = zd;
    return size;
  }
};

// BoundingBox::overlaps, BoundingBox.h:487-497 (closed intervals)
inline bool overlaps(const double* a, const double* b) {
  if (b[0] > a[3] || b[1] > a[4] || b[2] > a[5] || b[3] < a[0] || b[4] < a[1] || b[5] < a[2])
    return false;
  return true;
}

struct Contact {
  int b1, b2;
  V3 pos, normal;
  double dist, mu;
};



// ---------------------------------------------------------------------------------------------
// MaxContacts sphere/box/plane routines, pe/core/detection/fine/MaxContacts.h.
// ---------------------------------------------------------------------------------------------
struct Fine {
  orc_world* world;
  std::vector<Contact>* out;
  // Geometry ids: g < number of bodies names a body of the body storage, larger ids name union
  // members (g - number of bodies indexes orc_world::members).
  const Body& G(int g) const;
  int top(int g) const;  // RigidBody::getTopSuperBody
  void add(int b1, int b2, const V3& gpos, const V3& n, double dist);  // Contact.cpp:72-119
  void sphereSphere(int s1, int s2);
  void sphereBox(int s, int b);
  void spherePlane(int s, int p);
  void boxPlane(int b, int p);
  void boxBox(int b1, int b2);
  void sphereCapsule(int s, int c);
  void boxCapsule(int b, int c);
  void capsuleCapsule(int c1, int c2);
  void capsulePlane(int c, int p);
  void collide(int a, int b);
  void collideUnion(int a, int u);  // collide{Sphere,Box,Capsule,Plane}Union
};

// ---------------------------------------------------------------------------------------------
// Hierarchical hash grids with the reference's exact bookkeeping, HashGrids.h.
// ---------------------------------------------------------------------------------------------
struct HashGrid {
  struct Cell {
    std::vector<int>* bodies = nullptr;
    size_t occupiedCellsId = 0;
  };
  std::vector<Cell> cell;
  size_t xc, yc, zc, xm, ym, zm, xyc, xyzc, enlargementThreshold;
  double cellSpan, inverseCellSpan;
  std::vector<size_t> occupiedCells;
  size_t bodyCount = 0;

  // HashGrid::HashGrid, HashGrids.h (constructor)
  explicit HashGrid(double span) {
    xc = yc = zc = kCellCount;
    setup();
    cellSpan = span;
    inverseCellSpan = 1.0 / span;
  }
  ~HashGrid() { clear(); }
  void setup() {
    xm = xc - 1; ym = yc - 1; zm = zc - 1;
    xyc = xc * yc; xyzc = xyc * zc;
    enlargementThreshold = xyzc / kMinimalGridDensity;
    cell.assign(xyzc, Cell());
  }
  // Neighbour i of a cell, in the order of initializeNeighborOffsets (z outer, x inner),
  // with the periodic wrap of border cells.
  size_t neighbor(size_t c, unsigned i) const {
    const long x = (long)(c % xc), y = (long)((c / xc) % yc), z = (long)(c / xyc);
    const long dx = (long)(i % 3) - 1, dy = (long)((i / 3) % 3) - 1, dz = (long)(i / 9) - 1;
    const long nx = (x + dx + (long)xc) % (long)xc;
    const long ny = (y + dy + (long)yc) % (long)yc;
    const long nz = (z + dz + (long)zc) % (long)zc;
    return (size_t)nx + (size_t)ny * xc + (size_t)nz * xyc;
  }
  // HashGrid::hash, HashGrids.h:735-773
  size_t hash(const Body& b) const {
    const double x = b.aabb[0], y = b.aabb[1], z = b.aabb[2];
    size_t xh, yh, zh;
    if (x < 0) { double i = (-x) * inverseCellSpan; xh = xc - 1 - ((size_t)(i) & xm); }
    else       { double i = x * inverseCellSpan;    xh = (size_t)(i) & xm; }
    if (y < 0) { double i = (-y) * inverseCellSpan; yh = yc - 1 - ((size_t)(i) & ym); }
    else       { double i = y * inverseCellSpan;    yh = (size_t)(i) & ym; }
    if (z < 0) { double i = (-z) * inverseCellSpan; zh = zc - 1 - ((size_t)(i) & zm); }
    else       { double i = z * inverseCellSpan;    zh = (size_t)(i) & zm; }
    return xh + yh * xc + zh * xyc;
  }
  void add(std::vector<Body>& bodies, int b);       // HashGrids.h (HashGrid::add)
  void remove(std::vector<Body>& bodies, int b);    // HashGrid::remove
  void update(std::vector<Body>& bodies, int b);    // HashGrids.h:488-508
  void addToCell(std::vector<Body>& bodies, int b, size_t c);       // HashGrids.h:785-808
  void removeFromCell(std::vector<Body>& bodies, int b, size_t c);  // HashGrids.h:820-863
  void enlarge(std::vector<Body>& bodies);          // HashGrids.h:881-933
  void clear() {
    for (size_t c : occupiedCells) { delete cell[c].bodies; cell[c].bodies = nullptr; }
    occupiedCells.clear();
    bodyCount = 0;
  }
};

void HashGrid::addToCell(std::vector<Body>& bodies, int b, size_t c) {
  Cell& cl = cell[c];
  if (cl.bodies != nullptr) {
    bodies[b].cellId = cl.bodies->size();
    cl.bodies->push_back(b);
  } else {
    cl.bodies = new std::vector<int>();
    cl.bodies->reserve(16);
    bodies[b].cellId = 0;
    cl.bodies->push_back(b);
    cl.occupiedCellsId = occupiedCells.size();
    occupiedCells.push_back(c);
  }
}

void HashGrid::removeFromCell(std::vector<Body>& bodies, int b, size_t c) {
  Cell& cl = cell[c];
  if (cl.bodies->size() == 1) {
    delete cl.bodies;
    cl.bodies = nullptr;
    if (cl.occupiedCellsId == occupiedCells.size() - 1) {
      occupiedCells.pop_back();
    } else {
      const size_t last = occupiedCells.back();
      occupiedCells.pop_back();
      cell[last].occupiedCellsId = cl.occupiedCellsId;
      occupiedCells[cl.occupiedCellsId] = last;
    }
  } else {
    const size_t id = bodies[b].cellId;
    if (id == cl.bodies->size() - 1) {
      cl.bodies->pop_back();
    } else {
      const int lastEl = cl.bodies->back();
      cl.bodies->pop_back();
      bodies[lastEl].cellId = id;
      (*cl.bodies)[id] = lastEl;
    }
  }
}

void HashGrid::add(std::vector<Body>& bodies, int b) {
  if (bodyCount == enlargementThreshold) enlarge(bodies);
  const size_t h = hash(bodies[b]);
  bodies[b].hash = h;
  addToCell(bodies, b, h);
  ++bodyCount;
}

void HashGrid::remove(std::vector<Body>& bodies, int b) {
  removeFromCell(bodies, b, bodies[b].hash);
  --bodyCount;
}

void HashGrid::update(std::vector<Body>& bodies, int b) {
  const size_t nh = hash(bodies[b]);
  const size_t oh = bodies[b].hash;
  if (nh == oh) return;
  removeFromCell(bodies, b, oh);
  bodies[b].hash = nh;
  addToCell(bodies, b, nh);
}

void HashGrid::enlarge(std::vector<Body>& bodies) {
  std::vector<int> all;
  all.reserve(bodyCount);
  for (size_t c : occupiedCells)
    for (int e : *cell[c].bodies) all.push_back(e);
  clear();
  xc *= 2; yc *= 2; zc *= 2;
  setup();
  for (int e : all) add(bodies, e);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// The world: body storage, materials, detector, HCSITS state.
// ---------------------------------------------------------------------------------------------
struct orc_world {
  std::vector<Body> bodies;
  std::vector<Body> members;  // spheres contained in unions, grouped by union
  std::vector<Material> materials;
  std::vector<double> csf, cdf;  // pair tables, row-major n x n
  uint64_t nextSid = 1;
  V3 gravity;
  double damping = 1.0;
  // HCSITS defaults, HCSITS.h:408-431
  double erp = 0.7;
  size_t maxIterations = 100;
  size_t maxSubIterations = 20;
  double abortThreshold = 0.0;
  int relaxationModel = ORC_APPROX_INELASTIC_COULOMB_DECOUPLING;
  bool crTrig = false;

  // HashGrids state
  std::vector<int> bodiesToAdd, nonGridBodies;
  std::vector<HashGrid*> grids;        // grid objects, addressed by Body::grid
  std::list<int> gridList;             // ascending cell span, entries index grids
  bool gridActive = false;

  std::vector<Contact> contacts;
  std::vector<uint32_t> orderB1, orderB2, orderRank;  // orderRank empty: k-th occurrence order
  bool orderPending = false;
  // HCSITS caches, HCSITS.h:314-327
  std::vector<V3> v_, w_, dv_, dw_;
  std::vector<V3> r1_, r2_, n_, t_, o_, p_;
  std::vector<int> body1_, body2_;
  std::vector<double> dist_, mu_, diag_n_inv_;
  std::vector<M3> diag_nto_, diag_nto_inv_;
  std::vector<M2> diag_to_inv_;
  orc_stats stats{};

  ~orc_world() { for (HashGrid* g : grids) delete g; }

  double pairCsf(int a, int b) const { return csf[(size_t)a * materials.size() + b]; }
  double pairCdf(int a, int b) const { return cdf[(size_t)a * materials.size() + b]; }

  // ----- HashGrids top level -------------------------------------------------------------
  void hgAdd(int b) {  // HashGrids::add, HashGrids.h:961-968
    bodies[b].grid = -1;
    bodies[b].hash = bodiesToAdd.size();
    bodies[b].cellId = 0;
    bodiesToAdd.push_back(b);
  }
  int newGrid(double span) {
    grids.push_back(new HashGrid(span));
    return (int)grids.size() - 1;
  }
  // HashGrids::addGrid, HashGrids.h:1380-1446
  void addGrid(int b) {
    Body& body = bodies[b];
    const double size = body.finite() ? body.aabbSize() : -1.0;
    if (size > 0) {
      int grid = -1;
      if (gridList.empty()) {
        grid = newGrid(size * std::sqrt(kHierarchyFactor));
      } else {
        double cellSpan = 0;
        for (auto g = gridList.begin(); g != gridList.end(); ++g) {
          grid = *g;
          cellSpan = grids[grid]->cellSpan;
          if (size < cellSpan) {
            cellSpan /= kHierarchyFactor;
            if (size < cellSpan) {
              while (size < cellSpan) cellSpan /= kHierarchyFactor;
              grid = newGrid(cellSpan * kHierarchyFactor);
              gridList.insert(g, grid);
            }
            grids[grid]->add(bodies, b);
            bodies[b].grid = grid;
            return;
          }
        }
        while (size >= cellSpan) cellSpan *= kHierarchyFactor;
        grid = newGrid(cellSpan);
      }
      grids[grid]->add(bodies, b);
      bodies[b].grid = grid;
      gridList.push_back(grid);
      return;
    }
    body.grid = -1;
    body.hash = nonGridBodies.size();
    body.cellId = 1;
    nonGridBodies.push_back(b);
  }
  // HashGrids::addList, HashGrids.h:1469-1514
  void addList(int b) {
    if (nonGridBodies.size() == kGridActivationThreshold) {
      std::vector<int> tmp(nonGridBodies.begin(), nonGridBodies.end());
      nonGridBodies.clear();
      for (int e : tmp) addGrid(e);
      addGrid(b);
      gridActive = true;
      return;
    }
    bodies[b].grid = -1;
    bodies[b].hash = nonGridBodies.size();
    bodies[b].cellId = 1;
    nonGridBodies.push_back(b);
  }
  // HashGrids::findContacts, HashGrids.h:1121-1314
  void hgFindContacts(Fine& fine) {
    if (!bodiesToAdd.empty()) {
      for (int b : bodiesToAdd) {
        if (gridActive) addGrid(b); else addList(b);
      }
      bodiesToAdd.clear();
    }
    if (gridActive) {
      for (size_t b = 0; b < bodies.size(); ++b) {
        Body& body = bodies[b];
        if (body.grid >= 0) {
          HashGrid* grid = grids[body.grid];
          const double size = body.aabbSize();
          const double cellSpan = grid->cellSpan;
          if (size >= cellSpan || size < (cellSpan / kHierarchyFactor)) {
            grid->remove(bodies, (int)b);
            addGrid((int)b);
          } else {
            grid->update(bodies, (int)b);
          }
        }
      }
    }
    for (auto g = gridList.begin(); g != gridList.end(); ++g) {
      HashGrid* grid = grids[*g];
      // HashGrid::process, HashGrids.h:526-563
      std::vector<int> gridBodies;
      gridBodies.reserve(grid->bodyCount);
      for (size_t c : grid->occupiedCells) {
        const std::vector<int>& cb = *grid->cell[c].bodies;
        for (size_t a = 0; a < cb.size(); ++a) {
          for (size_t bb = a + 1; bb < cb.size(); ++bb) hgCollide(fine, cb[a], cb[bb]);
          gridBodies.push_back(cb[a]);
        }
        for (unsigned i = 0; i < 13; ++i) {
          const std::vector<int>* nb = grid->cell[grid->neighbor(c, i)].bodies;
          if (nb != nullptr) {
            for (int a : cb)
              for (int bb : *nb) hgCollide(fine, a, bb);
          }
        }
      }
      if (!gridBodies.empty()) {
        auto next = g;
        for (++next; next != gridList.end(); ++next) {
          // HashGrid::processBodies, HashGrids.h:583-607
          HashGrid* ng = grids[*next];
          for (int a : gridBodies) {
            const size_t c = ng->hash(bodies[a]);
            for (unsigned i = 0; i < 27; ++i) {
              const std::vector<int>* nb = ng->cell[ng->neighbor(c, i)].bodies;
              if (nb != nullptr)
                for (int bb : *nb) hgCollide(fine, a, bb);
            }
          }
        }
        for (int a : gridBodies)
          for (int bb : nonGridBodies) hgCollide(fine, a, bb);
      }
    }
    for (size_t a = 0; a < nonGridBodies.size(); ++a)
      for (size_t bb = a + 1; bb < nonGridBodies.size(); ++bb)
        hgCollide(fine, nonGridBodies[a], nonGridBodies[bb]);
  }
  // HashGrids::collide, HashGrids.h:1351-1358
  void hgCollide(Fine& fine, int a, int b) {
    if ((!bodies[a].fixed || !bodies[b].fixed) && overlaps(bodies[a].aabb, bodies[b].aabb))
      fine.collide(a, b);
  }
  // ExhaustiveSearch::findContacts, ExhaustiveSearch.h:212-232
  void exhaustiveFindContacts(Fine& fine) {
    for (size_t a = 0; a < bodies.size(); ++a)
      for (size_t b = a + 1; b < bodies.size(); ++b)
        hgCollide(fine, (int)a, (int)b);
  }

  // ----- HCSITS -------------------------------------------------------------------------
  void simulationStep(double dt);
  void resolveContacts(double dt);
  double relaxInelasticCoulombContactsByDecoupling(double dtinv);
  double relaxApproximateInelasticCoulombContactsByDecoupling(double dtinv);
  double relaxInelasticFrictionlessContacts(double dtinv);
  double relaxInelasticCoulombContactsByOrthogonalProjections(double dtinv, bool approximate);
  double relaxInelasticGeneralizedMaximumDissipationContacts(double dtinv);
  void initializeVelocityCorrections(Body& b, V3& dv, V3& dw, double dt);
  void integratePositions(Body& b, V3 v, V3 w, double dt);
  // Union::calcBoundingBox (UnionBase.cpp:208-230): the members' boxes merged in member order
  // (BoundingBox::operator+=, BoundingBox.h:339-348); every other body's own box.
  void calcBoundingBox(Body& b) {
    if (b.shape != ORC_UNION) { b.calcBoundingBox(); return; }
    for (int k = 0; k < b.mcount; ++k) {
      const double* a = members[b.mbegin + k].aabb;
      if (k == 0) { for (int j = 0; j < 6; ++j) b.aabb[j] = a[j]; continue; }
      for (int j = 0; j < 3; ++j) b.aabb[j] = (a[j] < b.aabb[j]) ? a[j] : b.aabb[j];
      for (int j = 3; j < 6; ++j) b.aabb[j] = (b.aabb[j] < a[j]) ? a[j] : b.aabb[j];
    }
  }
};

namespace {

// Contact::Contact, src/core/contact/Contact.cpp:72-119 (friction choice 102-105).
const Body& Fine::G(int g) const {
  const int nb = (int)world->bodies.size();
  return g < nb ? world->bodies[g] : world->members[g - nb];
}
int Fine::top(int g) const {
  const int nb = (int)world->bodies.size();
  return g < nb ? g : world->members[g - nb].super;
}

// The contact joins the top-level bodies (ContactBase.h:209-210) and takes the friction of the
// colliding geometries' materials (Contact.cpp:102-105).
void Fine::add(int g1, int g2, const V3& gpos, const V3& n, double dist) {
  const int b1 = top(g1), b2 = top(g2);
  const Body& A = G(b1);
  const Body& B = G(b2);
  const V3 rvel = A.velFromWF(gpos) - B.velFromWF(gpos);
  const double nvel = dot(n, rvel);
  const double tvel = dot(n, rvel - n * nvel);
  Contact c;
  c.b1 = b1; c.b2 = b2; c.pos = gpos; c.normal = n; c.dist = dist;
  const int m1 = G(g1).mat, m2 = G(g2).mat;
  c.mu = (std::fabs(tvel) > kFrictionThreshold) ? world->pairCdf(m1, m2) : world->pairCsf(m1, m2);
  out->push_back(c);
}

// MaxContacts::collideSphereSphere, MaxContacts.h:547-568
void Fine::sphereSphere(int s1, int s2) {
  if (G(s2).sid < G(s1).sid) std::swap(s1, s2);
  const Body& S1 = G(s1);
  const Body& S2 = G(s2);
  V3 normal = S1.gpos - S2.gpos;
  const double dist = length(normal) - S1.radius - S2.radius;
  if (dist < kContactThreshold) {
    normal = normalized(normal);
    const double k = S2.radius + 0.5 * dist;
    const V3 gPos = S2.gpos + normal * k;
    add(s1, s2, gPos, normal, dist);
  }
}

// MaxContacts::collideSphereBox, MaxContacts.h:589-667
void Fine::sphereBox(int s, int b) {
  const Body& S = G(s);
  const Body& B = G(b);
  const V3 l = 0.5 * B.lengths;
  const V3& spos = S.gpos;
  const V3& bpos = B.gpos;
  const M3& R = B.R;
  bool outside = false;
  const V3 d = spos - bpos;
  V3 p(d[0] * R[0] + d[1] * R[3] + d[2] * R[6], d[0] * R[1] + d[1] * R[4] + d[2] * R[7],
       d[0] * R[2] + d[1] * R[5] + d[2] * R[8]);
  if (p[0] < -l[0]) { p[0] = -l[0]; outside = true; }
  else if (p[0] > l[0]) { p[0] = l[0]; outside = true; }
  if (p[1] < -l[1]) { p[1] = -l[1]; outside = true; }
  else if (p[1] > l[1]) { p[1] = l[1]; outside = true; }
  if (p[2] < -l[2]) { p[2] = -l[2]; outside = true; }
  else if (p[2] > l[2]) { p[2] = l[2]; outside = true; }
  if (!outside) {
    double dist = std::fabs(p[0]) - l[0];
    int face = 0;
    for (int i = 1; i < 3; ++i) {
      const double tmp = std::fabs(p[i]) - l[i];
      if (dist < tmp) { dist = tmp; face = i; }
    }
    V3 n;
    n[face] = (p[face] > 0.0) ? 1.0 : -1.0;
    n = R * n;
    dist -= S.radius;
    add(s, b, spos, n, dist);
    return;
  }
  const V3 q = R * p;
  const V3 n = d - q;
  const double dist = length(n) - S.radius;
  if (dist < kContactThreshold) add(s, b, bpos + q, normalized(n), dist);
}

// MaxContacts::collideSpherePlane, MaxContacts.h:777-792
void Fine::spherePlane(int s, int p) {
  const Body& S = G(s);
  const Body& P = G(p);
  const double k = dot(P.normal, S.gpos);
  const double dist = k - S.radius - P.d;
  if (dist < kContactThreshold) {
    const V3 gPos = S.gpos - (S.radius + dist) * P.normal;
    add(s, p, gPos, P.normal, dist);
  }
}

// MaxContacts::collideBoxPlane, MaxContacts.h:2380-2528 (8 corners in the reference's order)
void Fine::boxPlane(int b, int p) {
  const Body& B = G(b);
  const Body& P = G(p);
  const V3& bpos = B.gpos;
  const V3& L = B.lengths;
  const M3& R = B.R;
  const double xr = 0.5 * L[0], yr = 0.5 * L[1], zr = 0.5 * L[2];
  const V3 nrel = trans(R) * P.normal;
  const double minlength = -0.99 * std::min(std::min(xr, yr), zr);
  static const int sgn[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int c = 0; c < 8; ++c) {
    const double sx = sgn[c][0] * xr, sy = sgn[c][1] * yr, sz = sgn[c][2] * zr;
    // (+-xr)*n0 (+-) yr*n1 (+-) zr*n2, evaluated left to right
    double t = sx * nrel[0];
    t = (sgn[c][1] < 0) ? t - yr * nrel[1] : t + yr * nrel[1];
    t = (sgn[c][2] < 0) ? t - zr * nrel[2] : t + zr * nrel[2];
    if (t < minlength) {
      const V3 pos = bpos + R * V3(sx, sy, sz);
      const double k = dot(pos, P.normal);
      const double dist = k - P.d;
      if (dist < kContactThreshold) add(b, p, pos, P.normal, dist);
    }
  }
}

// intersectLines, src/core/GeomTools.cpp:387-408
void intersectLines(const V3& o1, const V3& d1, const V3& o2, const V3& d2, double& s, double& t) {
  const double sqrlen = dot(cross(d1, d2), cross(d1, d2));
  if (sqrlen < 1e-8) {  // parallelThreshold, Thresholds.h:170
    s = 0.0;
    t = 0.0;
  } else {
    const double isqrlen = 1.0 / sqrlen;
    const V3 p = o2 - o1;
    const double dt = dot(d1, d2);
    const double a = dot(d1, p);
    const double b = -dot(d2, p);
    s = (a * dot(d2, d2) + dt * b) * isqrlen;
    t = (b * dot(d1, d1) + dt * a) * isqrlen;
  }
}

// MaxContacts::collideBoxBox, MaxContacts.h:867-2268: separating-axis test over 3 + 3 face
// axes and 9 edge axes, then one edge-edge contact or a clipped face contact (up to 8).
void Fine::boxBox(int i1, int i2) {
  if (G(i2).sid < G(i1).sid) std::swap(i1, i2);
  const Body* B1 = &G(i1);
  const Body* B2 = &G(i2);
  double maxDepth = -kInf;
  bool invertNormal = false;
  unsigned contactCase = 0;
  V3 contactNormal, normal_c;
  const M3& R1 = B1->R;
  const M3& R2 = B2->R;
  const M3 R = trans(R1) * R2;  // b2_rR
  M3 Q;                          // b2_rQ = fabs(b2_rR)
  for (int k = 0; k < 9; ++k) Q[k] = std::fabs(R[k]);
  const V3 P = trans(R1) * (B2->gpos - B1->gpos);  // b2_rPos
  const double hl1[3] = {0.5 * B1->lengths[0], 0.5 * B1->lengths[1], 0.5 * B1->lengths[2]};
  const double hl2[3] = {0.5 * B2->lengths[0], 0.5 * B2->lengths[1], 0.5 * B2->lengths[2]};
  // face axes of box 1 (cases 1-3)
  for (int a = 0; a < 3; ++a) {
    const double term1 = std::fabs(P[a]) - (hl1[a] + hl2[0] * Q[3 * a] + hl2[1] * Q[3 * a + 1] + hl2[2] * Q[3 * a + 2]);
    if (term1 > kContactThreshold) return;
    if (a == 0 || term1 > maxDepth) {
      maxDepth = term1;
      contactNormal = V3(R1[a], R1[3 + a], R1[6 + a]);
      contactCase = 1 + a;
      invertNormal = (P[a] < 0.0);
    }
  }
  // face axes of box 2 (cases 4-6)
  for (int a = 0; a < 3; ++a) {
    const double term1 = P[0] * R[a] + P[1] * R[3 + a] + P[2] * R[6 + a];
    const double term2 = std::fabs(term1) - (hl1[0] * Q[a] + hl1[1] * Q[3 + a] + hl1[2] * Q[6 + a] + hl2[a]);
    if (term2 > kContactThreshold) return;
    if (term2 > maxDepth) {
      maxDepth = term2;
      contactNormal = V3(R2[a], R2[3 + a], R2[6 + a]);
      contactCase = 4 + a;
      invertNormal = (term1 < 0.0);
    }
  }
  // edge axes (cases 7-15): term1 = P[pa]*R[ra] - P[pb]*R[rb]; term2 = hl1[h]*Q[q] + hl1[h']*Q[q'];
  // term3 = hl2[h]*Q[q] + hl2[h']*Q[q']; length = sqrt(R[la]^2 + R[lb]^2); normal_c components
  // as (sign, index) with sign 0 for a zero component.
  static const int E[9][14] = {
      {2, 3, 1, 6, 1, 6, 2, 3, 2, 1, 1, 2, 6, 3},  // case 7
      {2, 4, 1, 7, 1, 7, 2, 4, 0, 2, 2, 0, 7, 4},  // case 8
      {2, 5, 1, 8, 1, 8, 2, 5, 1, 0, 0, 1, 8, 5},  // case 9
      {0, 6, 2, 0, 2, 0, 0, 6, 2, 4, 1, 5, 6, 0},  // case 10
      {0, 7, 2, 1, 2, 1, 0, 7, 0, 5, 2, 3, 7, 1},  // case 11
      {0, 8, 2, 2, 2, 2, 0, 8, 1, 3, 0, 4, 8, 2},  // case 12
      {1, 0, 0, 3, 0, 3, 1, 0, 2, 7, 1, 8, 3, 0},  // case 13
      {1, 1, 0, 4, 0, 4, 1, 1, 0, 8, 2, 6, 4, 1},  // case 14
      {1, 2, 0, 5, 0, 5, 1, 2, 1, 6, 0, 7, 5, 2}}; // case 15
  static const int NC[9][6] = {{0, 0, -1, 6, 1, 3}, {0, 0, -1, 7, 1, 4}, {0, 0, -1, 8, 1, 5},
                               {1, 6, 0, 0, -1, 0}, {1, 7, 0, 0, -1, 1}, {1, 8, 0, 0, -1, 2},
                               {-1, 3, 1, 0, 0, 0}, {-1, 4, 1, 1, 0, 0}, {-1, 5, 1, 2, 0, 0}};
  for (int e = 0; e < 9; ++e) {
    const int* t = E[e];
    const double term1 = P[t[0]] * R[t[1]] - P[t[2]] * R[t[3]];
    const double term2 = hl1[t[4]] * Q[t[5]] + hl1[t[6]] * Q[t[7]];
    const double term3 = hl2[t[8]] * Q[t[9]] + hl2[t[10]] * Q[t[11]];
    double sum = std::fabs(term1) - (term2 + term3);
    if (std::fabs(term1) > term2 + term3 + kContactThreshold) return;
    const double length = std::sqrt(R[t[12]] * R[t[12]] + R[t[13]] * R[t[13]]);
    if (length > std::numeric_limits<double>::epsilon()) {
      sum /= length;
      if (sum > maxDepth && std::fabs(sum - maxDepth) > kAccuracy) {
        maxDepth = sum;
        double c[3];
        for (int k = 0; k < 3; ++k) {
          const int sg = NC[e][2 * k], ix = NC[e][2 * k + 1];
          c[k] = sg == 0 ? 0.0 : (sg > 0 ? R[ix] / length : -R[ix] / length);
        }
        normal_c = V3(c[0], c[1], c[2]);
        contactCase = 7 + e;
        invertNormal = (term1 < 0.0);
      }
    }
  }
  if (contactCase == 0) return;
  if (contactCase > 6) contactNormal = R1 * normal_c;
  if (invertNormal) contactNormal = -contactNormal;

  if (contactCase > 6) {  // edge-edge
    V3 pB1 = B1->gpos;
    V3 sign = trans(R1) * contactNormal;
    for (int i = 0; i < 3; ++i) sign[i] = (sign[i] > 0) ? hl1[i] : -hl1[i];
    pB1 += R1 * sign;
    V3 pB2 = B2->gpos;
    sign = trans(R2) * contactNormal;
    for (int i = 0; i < 3; ++i) sign[i] = (sign[i] > 0) ? -hl2[i] : hl2[i];
    pB2 += R2 * sign;
    V3 ua, ub;
    for (int i = 0; i < 3; ++i) {
      ua[i] = R1[(contactCase - 7) / 3 + i * 3];
      ub[i] = R2[(contactCase - 7) % 3 + i * 3];
    }
    double s, t;
    intersectLines(pB1, ua, pB2, ub, s, t);
    pB1 += s * ua;
    pB2 += t * ub;
    const V3 gpos = 0.5 * (pB1 + pB2);
    add(i2, i1, gpos, contactNormal, maxDepth);  // addEdgeEdgeContact(b2, b1, ...); edges unused by HCSITS
    return;
  }

  // face case: the reference face belongs to box "a", the incident face to box "b"
  const double* hla = hl1;
  const double* hlb = hl2;
  if (contactCase > 3) {
    std::swap(i1, i2);
    std::swap(B1, B2);
    std::swap(hla, hlb);
    contactNormal = -contactNormal;
  }
  const M3& Ra = B1->R;
  const M3& Rb = B2->R;
  const V3 normala = trans(Ra) * contactNormal;
  const V3 normalb = trans(Rb) * contactNormal;
  unsigned xb, yb, zb;
  if (std::fabs(normalb[0]) > std::fabs(normalb[1])) {
    if (std::fabs(normalb[0]) > std::fabs(normalb[2])) { xb = 0; yb = 1; zb = 2; }
    else { xb = 2; yb = 0; zb = 1; }
  } else if (std::fabs(normalb[1]) > std::fabs(normalb[2])) {
    xb = 1; yb = 2; zb = 0;
  } else {
    xb = 2; yb = 0; zb = 1;
  }
  const unsigned xa = (contactCase < 4) ? contactCase - 1 : contactCase - 4;
  const unsigned ya = (xa == 0) ? 1 : (xa == 1 ? 2 : 0);
  const unsigned za = (xa == 0) ? 2 : (xa == 1 ? 0 : 1);
  V3 vertexb[4];
  for (int k = 0; k < 4; ++k) {
    vertexb[k][xb] = (normalb[xb] > 0.0) ? -hlb[xb] : hlb[xb];
    vertexb[k][yb] = (k & 1) ? hlb[yb] : -hlb[yb];
    vertexb[k][zb] = (k & 2) ? hlb[zb] : -hlb[zb];
  }
  const V3 ab = B2->gpos - B1->gpos;
  V3 vba[4];
  for (int k = 0; k < 4; ++k) vba[k] = trans(Ra) * (ab + Rb * vertexb[k]);
  const double offseta = (normala[xa] > 0.0) ? hla[xa] : -hla[xa];
  // incident-face edges (0,1) (0,2) (3,1) (3,2), each clipped at ya = +-hla[ya], za = +-hla[za]
  static const int edges[4][2] = {{0, 1}, {0, 2}, {3, 1}, {3, 2}};
  for (int e = 0; e < 4; ++e) {
    const V3& v0 = vba[edges[e][0]];
    const V3& v1 = vba[edges[e][1]];
    for (int side = 0; side < 4; ++side) {
      const unsigned d = (side < 2) ? ya : za;
      const unsigned o = (side < 2) ? za : ya;
      const double bound = (side & 1) ? -hla[d] : hla[d];
      if ((v0[d] > bound) ^ (v1[d] > bound)) {
        const double s = (bound - v0[d]) / (v1[d] - v0[d]);
        double tmp;
        if (s > 0.0 && s < 1.0 && std::fabs(tmp = v0[o] + s * (v1[o] - v0[o])) < hla[o]) {
          const double dist = std::fabs(v0[xa] + s * (v1[xa] - v0[xa])) - hla[xa];
          if (dist < kContactThreshold) {
            V3 posa;
            posa[xa] = offseta;
            posa[d] = bound;
            posa[o] = tmp;
            add(i2, i1, B1->gpos + Ra * posa, contactNormal, dist);
          }
        }
      }
    }
  }
  for (int k = 0; k < 4; ++k) {  // incident vertices inside the reference face
    double dist;
    if (std::fabs(vba[k][ya]) <= hla[ya] && std::fabs(vba[k][za]) <= hla[za] &&
        (dist = std::fabs(vba[k][xa]) - hla[xa]) < kContactThreshold)
      add(i2, i1, B1->gpos + Ra * vba[k], contactNormal, dist);
  }
  V3 vertexa[4];
  for (int k = 0; k < 4; ++k) {
    vertexa[k][xa] = (normala[xa] > 0.0) ? hla[xa] : -hla[xa];
    vertexa[k][ya] = (k & 1) ? hla[ya] : -hla[ya];
    vertexa[k][za] = (k & 2) ? hla[za] : -hla[za];
  }
  const V3 ba = B1->gpos - B2->gpos;
  for (int k = 0; k < 4; ++k) {  // reference corners inside the incident face
    const V3 vab = trans(Rb) * (ba + Ra * vertexa[k]);
    double dist;
    if (((std::fabs(vab[yb]) <= hlb[yb] && std::fabs(vab[zb]) < hlb[zb]) ||
         (std::fabs(vab[yb]) < hlb[yb] && std::fabs(vab[zb]) <= hlb[zb])) &&
        (dist = std::fabs(vab[xb]) - hlb[xb]) < kContactThreshold)
      add(i2, i1, B2->gpos + Rb * vab, contactNormal, dist);
  }
}

// getClosestLineSegmentPoints, src/core/GeomTools.cpp:259-385: the closest points of the
// segments [a1, a2] and [b1, b2]; the end-point cases first, then the end point against the
// other segment's interior, then the two interiors.
void closestSegmentPoints(const V3& a1, const V3& a2, const V3& b1, const V3& b2, V3& cp1, V3& cp2) {
  const V3 a1a2 = a2 - a1, b1b2 = b2 - b1, a1b1 = b1 - a1;
  const double da1 = dot(a1a2, a1b1), db1 = dot(b1b2, a1b1);
  if (da1 <= 0.0 && db1 >= 0.0) { cp1 = a1; cp2 = b1; return; }
  const V3 a1b2 = b2 - a1;
  const double da2 = dot(a1a2, a1b2), db2 = dot(b1b2, a1b2);
  if (da2 <= 0.0 && db2 <= 0.0) { cp1 = a1; cp2 = b2; return; }
  const V3 a2b1 = b1 - a2;
  const double da3 = dot(a1a2, a2b1), db3 = dot(b1b2, a2b1);
  if (da3 >= 0.0 && db3 >= 0.0) { cp1 = a2; cp2 = b1; return; }
  const V3 a2b2 = b2 - a2;
  const double da4 = dot(a1a2, a2b2), db4 = dot(b1b2, a2b2);
  if (da4 >= 0.0 && db4 <= 0.0) { cp1 = a2; cp2 = b2; return; }
  const double la = dot(a1a2, a1a2);
  if (da1 >= 0.0 && da3 <= 0.0) {
    const V3 k = (da1 / la) * a1a2;
    if (dot(b1b2, a1b1 - k) >= 0.0) { cp1 = a1 + k; cp2 = b1; return; }
  }
  if (da2 >= 0.0 && da4 <= 0.0) {
    const V3 k = (da2 / la) * a1a2;
    if (dot(b1b2, a1b2 - k) <= 0.0) { cp1 = a1 + k; cp2 = b2; return; }
  }
  const double lb = dot(b1b2, b1b2);
  if (db1 <= 0.0 && db2 >= 0.0) {
    const V3 k = (-db1 / lb) * b1b2;
    if (dot(a1a2, -a1b1 - k) >= 0.0) { cp1 = a1; cp2 = b1 + k; return; }
  }
  if (db3 <= 0.0 && db4 >= 0.0) {
    const V3 k = (-db3 / lb) * b1b2;
    if (dot(a1a2, -a2b1 - k) <= 0.0) { cp1 = a2; cp2 = b1 + k; return; }
  }
  const double scale = dot(a1a2, b1b2);
  double div = la * lb - scale * scale;
  div = 1.0 / div;
  const double sa = (lb * da1 - scale * db1) * div;
  const double tb = (scale * da1 - la * db1) * div;
  cp1 = a1 + sa * a1a2;
  cp2 = b1 + tb * b1b2;
}

// getClosestLineBoxPoints, src/core/GeomTools.cpp:73-257: the closest points of the segment
// [p1, p2] and the box (centre c, rotation R, side lengths `side`).  In the box frame, with the
// segment's direction mirrored into the positive octant, the squared distance is a convex
// piecewise quadratic in the segment parameter t whose pieces change where the mirrored point
// crosses a face plane (tanchor); the walk follows the pieces from t = 0 until the derivative
// changes sign, then clamps the segment point to the box.
void closestLineBoxPoints(const V3& p1, const V3& p2, const V3& c, const M3& R, const V3& side, V3& lret,
                          V3& bret) {
  const double eps = std::numeric_limits<double>::epsilon();
  const V3 l = p2 - p1;
  V3 tmp = trans(R) * l;
  const double sgn[3] = {tmp[0] < 0 ? -1.0 : 1.0, tmp[1] < 0 ? -1.0 : 1.0, tmp[2] < 0 ? -1.0 : 1.0};
  const double v[3] = {sgn[0] * tmp[0], sgn[1] * tmp[1], sgn[2] * tmp[2]};
  const double v2[3] = {v[0] * v[0], v[1] * v[1], v[2] * v[2]};
  tmp = p1 - c;
  const double s[3] = {sgn[0] * (R[0] * tmp[0] + R[3] * tmp[1] + R[6] * tmp[2]),
                       sgn[1] * (R[1] * tmp[0] + R[4] * tmp[1] + R[7] * tmp[2]),
                       sgn[2] * (R[2] * tmp[0] + R[5] * tmp[1] + R[8] * tmp[2])};
  const double h[3] = {0.5 * side[0], 0.5 * side[1], 0.5 * side[2]};
  int region[3] = {0, 0, 0};
  double tanchor[3] = {2, 2, 2};
  double dd2dt = 0;
  for (int i = 0; i < 3; ++i) {
    if (v[i] > eps) {
      if (s[i] < -h[i]) {
        region[i] = -1;
        tanchor[i] = (-h[i] - s[i]) / v[i];
        dd2dt -= v2[i] * tanchor[i];
      } else if (s[i] > h[i]) {
        region[i] = 1;
        tanchor[i] = (h[i] - s[i]) / v[i];
        dd2dt -= v2[i] * tanchor[i];
      } else {
        tanchor[i] = (h[i] - s[i]) / v[i];
      }
    }
  }
  double t = 0;
  if (dd2dt < -kAccuracy) {
    do {
      double next_t = 1;
      for (int i = 0; i < 3; ++i)
        if (tanchor[i] > t && tanchor[i] < 1.0 && tanchor[i] < next_t) next_t = tanchor[i];
      double next_dd2dt = 0;
      for (int i = 0; i < 3; ++i)
        if (region[i] != 0) next_dd2dt += v2[i] * (next_t - tanchor[i]);
      if (next_dd2dt >= 0) {
        t -= dd2dt / ((next_dd2dt - dd2dt) / (next_t - t));
        break;
      }
      for (int i = 0; i < 3; ++i)
        if (tanchor[i] == next_t) {
          tanchor[i] = (h[i] - s[i]) / v[i];
          ++region[i];
        }
      t = next_t;
      dd2dt = next_dd2dt;
    } while (t < 1.0);
  }
  lret = p1 + t * l;
  for (int i = 0; i < 3; ++i) {
    double x = sgn[i] * (s[i] + t * v[i]);
    if (x < -h[i]) x = -h[i];
    else if (x > h[i]) x = h[i];
    tmp[i] = x;
  }
  bret = (R * tmp) + c;
}

// Vector3 operator==, Vector3.h:1421-1428: every component within 1e-8 (Equal.h:172-175)
bool nearly_equal(const V3& a, const V3& b) {
  return std::fabs(a[0] - b[0]) < 1e-8 && std::fabs(a[1] - b[1]) < 1e-8 && std::fabs(a[2] - b[2]) < 1e-8;
}

// collideSphereCapsule, MaxContacts.h:690-725: the sphere against the closest point of the
// capsule's axis segment.  The sphere is body1.
void Fine::sphereCapsule(int si, int ci) {
  const Body& S = G(si);
  const Body& C = G(ci);
  const double hl = 0.5 * C.length;
  const V3& spos = S.gpos;
  const V3& cpos = C.gpos;
  const M3& R = C.R;
  double sx = R[0] * (spos[0] - cpos[0]) + R[3] * (spos[1] - cpos[1]) + R[6] * (spos[2] - cpos[2]);
  if (sx > hl) sx = hl;
  else if (sx < -hl) sx = -hl;
  const V3 spos2(sx * R[0] + cpos[0], sx * R[3] + cpos[1], sx * R[6] + cpos[2]);
  V3 normal = spos - spos2;
  const double dist = length(normal) - S.radius - C.radius;
  if (dist < kContactThreshold) {
    normal = normalized(normal);
    const double k = C.radius + 0.5 * dist;
    add(si, ci, spos2 + normal * k, normal, dist);
  }
}

// collideBoxCapsule, MaxContacts.h:2273-2330: the closest points of the axis segment and the
// box, from each end of the segment; the second contact is added whenever its points differ
// from the first's (without a distance test, as in the reference).  The box is body1.
void Fine::boxCapsule(int bi, int ci) {
  const Body& B = G(bi);
  const Body& C = G(ci);
  const M3& R = C.R;
  const V3 c_up(0.5 * C.length * R[0], 0.5 * C.length * R[3], 0.5 * C.length * R[6]);
  const V3 up = C.gpos + c_up, down = C.gpos - c_up;
  V3 cp1, bp1;
  closestLineBoxPoints(up, down, B.gpos, B.R, B.lengths, cp1, bp1);
  V3 normal1 = bp1 - cp1;
  const double dist1 = length(normal1) - C.radius;
  if (dist1 > kContactThreshold) return;
  normal1 = normalized(normal1);
  add(bi, ci, bp1 - 0.5 * dist1 * normal1, normal1, dist1);
  V3 cp2, bp2;
  closestLineBoxPoints(down, up, B.gpos, B.R, B.lengths, cp2, bp2);
  if (nearly_equal(cp1, cp2)) return;
  V3 normal2 = bp2 - cp2;
  const double dist2 = length(normal2) - C.radius;
  normal2 = normalized(normal2);
  add(bi, ci, bp2 - 0.5 * dist2 * normal2, normal2, dist2);
}

// collideCapsuleCapsule, MaxContacts.h:2603-2690: body1 is the capsule with the smaller SID.
// Nearly parallel, overlapping axes give two contacts at the ends of the overlap; otherwise
// one contact between the axes' closest points.
void Fine::capsuleCapsule(int c1i, int c2i) {
  if (G(c2i).sid < G(c1i).sid) std::swap(c1i, c2i);
  const Body& C1 = G(c1i);
  const Body& C2 = G(c2i);
  const V3 d1(C1.R[0], C1.R[3], C1.R[6]), d2(C2.R[0], C2.R[3], C2.R[6]);
  const V3 up1 = (0.5 * C1.length) * d1, up2 = (0.5 * C2.length) * d2;
  V3 cp1, cp2;
  closestSegmentPoints(C1.gpos + up1, C1.gpos - up1, C2.gpos + up2, C2.gpos - up2, cp1, cp2);
  V3 normal = cp1 - cp2;
  const double dist = length(normal) - C1.radius - C2.radius;
  if (!(dist < kContactThreshold)) return;
  normal = normalized(normal);
  const double c2x = dot(d1, C2.gpos - C1.gpos);
  if ((std::fabs(dot(d1, d2)) - 1.0) > -kParallelThreshold && c2x > -C1.length && c2x < C1.length) {
    const double k = C1.radius + 0.5 * dist;
    const double hl1 = 0.5 * C1.length, hl2 = 0.5 * C2.length;
    if (hl1 < c2x + hl2) add(c1i, c2i, C1.gpos + up1 - k * normal, normal, dist);
    else add(c1i, c2i, C2.gpos + hl2 * d1 + k * normal, normal, dist);
    if (-hl1 > c2x - hl2) add(c1i, c2i, C1.gpos - up1 - k * normal, normal, dist);
    else add(c1i, c2i, C2.gpos - hl2 * d1 + k * normal, normal, dist);
  } else {
    const double k = C2.radius + 0.5 * dist;
    add(c1i, c2i, cp2 + normal * k, normal, dist);
  }
}

// collideCapsulePlane, MaxContacts.h:2744-2780: each end sphere against the plane.  The
// capsule is body1.
void Fine::capsulePlane(int ci, int pi) {
  const Body& C = G(ci);
  const Body& P = G(pi);
  const V3 c_up = (0.5 * C.length) * V3(C.R[0], C.R[3], C.R[6]);
  const V3 posUp = C.gpos + c_up, posDn = C.gpos - c_up;
  const double distUp = dot(posUp, P.normal) - P.d - C.radius;
  const double distDn = dot(posDn, P.normal) - P.d - C.radius;
  if (distUp < kContactThreshold) add(ci, pi, posUp - C.radius * P.normal, P.normal, distUp);
  if (distDn < kContactThreshold) add(ci, pi, posDn - C.radius * P.normal, P.normal, distDn);
}

// MaxContacts::collide dispatch, MaxContacts.h:242-526 (sphere/box/capsule/plane subset)
// collideSphereUnion, collideBoxUnion, collideCapsuleUnion, collidePlaneUnion (MaxContacts.h:843-848,
// 2576-2581, 2828-2833, 3453-3457): the other body against every member, in the union's order.
void Fine::collideUnion(int a, int u) {
  const Body& U = G(u);
  const int nb = (int)world->bodies.size();
  for (int k = 0; k < U.mcount; ++k) {
    const int m = nb + U.mbegin + k;
    if (G(a).shape == ORC_PLANE) collide(m, a);
    else collide(a, m);
  }
}

void Fine::collide(int a, int b) {
  const int ta = G(a).shape, tb = G(b).shape;
  if (ta == ORC_UNION && tb == ORC_UNION) {  // collideUnionUnion, MaxContacts.h:3569-3581
    if (G(b).sid < G(a).sid) std::swap(a, b);
    const Body& U1 = G(a);
    const Body& U2 = G(b);
    const int nb = (int)world->bodies.size();
    for (int k1 = 0; k1 < U1.mcount; ++k1)
      for (int k2 = 0; k2 < U2.mcount; ++k2) collide(nb + U1.mbegin + k1, nb + U2.mbegin + k2);
  } else if (tb == ORC_UNION) {
    collideUnion(a, b);
  } else if (ta == ORC_UNION) {
    collideUnion(b, a);
  } else if (ta == ORC_SPHERE) {
    if (tb == ORC_SPHERE) sphereSphere(a, b);
    else if (tb == ORC_BOX) sphereBox(a, b);
    else if (tb == ORC_CAPSULE) sphereCapsule(a, b);
    else spherePlane(a, b);
  } else if (ta == ORC_BOX) {
    if (tb == ORC_SPHERE) sphereBox(b, a);
    else if (tb == ORC_BOX) boxBox(a, b);
    else if (tb == ORC_CAPSULE) boxCapsule(a, b);
    else boxPlane(a, b);
  } else if (ta == ORC_CAPSULE) {
    if (tb == ORC_SPHERE) sphereCapsule(b, a);
    else if (tb == ORC_BOX) boxCapsule(b, a);
    else if (tb == ORC_CAPSULE) capsuleCapsule(a, b);
    else capsulePlane(a, b);
  } else if (ta == ORC_PLANE) {
    if (tb == ORC_SPHERE) spherePlane(b, a);
    else if (tb == ORC_BOX) boxPlane(b, a);
    else if (tb == ORC_CAPSULE) capsulePlane(b, a);
    // plane-plane: no contact
  }
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// CollisionSystem<HCSITS>::simulationStep, HCSITS.h:1472-1531
void orc_world::simulationStep(double dt) {
  double t0 = now();
  contacts.clear();
  Fine fine{this, &contacts};
  hgFindContacts(fine);
  stats.t_detection = now() - t0;
  if (orderPending) {
    orderPending = false;
    if (orderB1.size() != contacts.size()) throw std::runtime_error("sweep order does not match the contacts");
    std::vector<Contact> sorted;
    std::vector<bool> taken(contacts.size(), false);
    for (size_t k = 0; k < orderB1.size(); ++k) {
      size_t found = contacts.size();
      uint32_t seen = 0;
      for (size_t c = 0; c < contacts.size(); ++c) {
        if ((uint32_t)contacts[c].b1 != orderB1[k] || (uint32_t)contacts[c].b2 != orderB2[k]) continue;
        if (orderRank.empty() ? !taken[c] : seen++ == orderRank[k]) { found = c; break; }
      }
      if (found < contacts.size() && taken[found]) throw std::runtime_error("sweep order repeats a contact");
      if (found == contacts.size()) throw std::runtime_error("sweep order does not match the contacts");
      taken[found] = true;
      sorted.push_back(contacts[found]);
    }
    contacts.swap(sorted);
  }
  resolveContacts(dt);
}

// HCSITS.h:4094-4109
void orc_world::initializeVelocityCorrections(Body& b, V3& dv, V3& dw, double dt) {
  if (b.awake) {
    if (!b.fixed) {
      dv = (b.invMass * dt) * b.force;
      dw = dt * (b.invInertiaWF() * b.torque);
    }
  }
  b.force = V3();  // Settings::forceReset() is true (Settings.cpp:46)
  b.torque = V3();
}

// HCSITS.h:4135-4184
void orc_world::integratePositions(Body& b, V3 v, V3 w, double dt) {
  if (b.awake) {
    b.gpos += v * dt;
    const V3 phi = w * dt;
    const Q dq = quat_axis_angle(phi, length(phi), crTrig);
    b.q = quat_mul(dq, b.q);
    b.R = quat_to_R(b.q);
    if (damping < 1.0) {
      const double drag = std::pow(damping, dt);
      v = v * drag;
      w = w * drag;
    }
    b.v = v;
    b.w = w;
    if (b.shape == ORC_UNION) {
      // SuperBody::updateBody -> Sphere::update(dq), Sphere.cpp:1052-1070: each member at the
      // union's position plus its rotated relative position
      for (int k = 0; k < b.mcount; ++k) {
        Body& m = members[b.mbegin + k];
        m.gpos = b.gpos + b.R * m.rpos;
        m.q = quat_mul(dq, m.q);
        m.R = quat_to_R(m.q);
        m.calcBoundingBox();
      }
    } else if (b.shape == ORC_PLANE) {
      b.normal = quat_rotate(b.q, V3(0, 0, 1));
      b.d = dot(b.normal, b.gpos);
    }
    calcBoundingBox(b);
    // calcMotion never puts a body to sleep: sleepThreshold = 0 (pe/config/SleepMode.h:37)
  }
}

// CollisionSystem<HCSITS>::resolveContacts, HCSITS.h:1552-2019 (serial: every contact is kept)
void orc_world::resolveContacts(double dt) {
  const double dtinv = 1.0 / dt;
  double t0 = now();
  const size_t nc = contacts.size();
  body1_.resize(nc); body2_.resize(nc);
  r1_.resize(nc); r2_.resize(nc); n_.resize(nc); t_.resize(nc); o_.resize(nc);
  dist_.resize(nc); mu_.resize(nc);
  diag_nto_.resize(nc); diag_nto_inv_.resize(nc); diag_to_inv_.resize(nc); diag_n_inv_.resize(nc);
  p_.resize(nc);
  stats.max_penetration = 0;
  stats.num_contacts = nc;
  // Contact caching, HCSITS.h:1741-1818
  for (size_t j = 0; j < nc; ++j) {
    const Contact& c = contacts[j];
    const Body& b1 = bodies[c.b1];
    const Body& b2 = bodies[c.b2];
    body1_[j] = c.b1;
    body2_[j] = c.b2;
    r1_[j] = c.pos - b1.gpos;
    r2_[j] = c.pos - b2.gpos;
    n_[j] = c.normal;
    const V3& n = n_[j];
    if (std::fabs(n[0]) < std::fabs(n[1])) {
      if (std::fabs(n[0]) < std::fabs(n[2])) t_[j] = V3(0, n[2], -n[1]);
      else t_[j] = V3(n[1], -n[0], 0);
    } else {
      if (std::fabs(n[1]) < std::fabs(n[2])) t_[j] = V3(-n[2], 0, n[0]);
      else t_[j] = V3(n[1], -n[0], 0);
    }
    t_[j] = normalized(t_[j]);
    o_[j] = cross(n_[j], t_[j]);
    const M3 cf = M3::cols(n_[j], t_[j], o_[j]);
    dist_[j] = c.dist;
    if (dist_[j] < 0) {
      stats.max_penetration = std::max(stats.max_penetration, -dist_[j]);
      dist_[j] *= erp;
    }
    mu_[j] = c.mu;
    M3 diag = -(mul_skew(skew_mul(r1_[j], b1.invInertiaWF()), r1_[j]) +
                mul_skew(skew_mul(r2_[j], b2.invInertiaWF()), r2_[j]));
    diag[0] += b1.invMass + b2.invMass;
    diag[4] += b1.invMass + b2.invMass;
    diag[8] += b1.invMass + b2.invMass;
    diag = trans(cf) * diag * cf;
    diag_nto_[j] = diag;
    diag_nto_inv_[j] = inverse(diag);
    diag_n_inv_[j] = 1.0 / diag[0];
    M2 to; to.m[0] = diag[4]; to.m[1] = diag[5]; to.m[2] = diag[7]; to.m[3] = diag[8];
    diag_to_inv_[j] = inverse2(to);
    p_[j] = V3();
  }
  stats.t_contact_caching = now() - t0;
  t0 = now();
  // Body caching, HCSITS.h:1834-1878
  const size_t nb = bodies.size();
  v_.resize(nb); w_.resize(nb); dv_.resize(nb); dw_.resize(nb);
  for (size_t j = 0; j < nb; ++j) {
    Body& b = bodies[j];
    b.awake = true;
    initializeVelocityCorrections(b, dv_[j], dw_[j], dt);
    if (b.awake && !b.fixed) {
      v_[j] = b.v + gravity * dt;
#if ORC_PLANT == 1
      w_[j] = b.w + dt * (b.invInertiaWF() * cross(b.w, b.inertiaWF() * b.w));
#elif ORC_PLANT == 3
      w_[j] = b.w;
#else
      w_[j] = b.w + dt * (b.invInertiaWF() * cross(b.inertiaWF() * b.w, b.w));
#endif
    } else {
      v_[j] = b.v;
      w_[j] = b.w;
    }
  }
  stats.t_body_caching = now() - t0;
  t0 = now();
  // Sweep loop, HCSITS.h:1891-1947
  size_t it = 0;
  double delta = 0;
  for (it = 0; it < maxIterations; ++it) {
    switch (relaxationModel) {
      case ORC_INELASTIC_FRICTIONLESS: delta = relaxInelasticFrictionlessContacts(dtinv); break;
      case ORC_APPROX_INELASTIC_COULOMB_DECOUPLING:
        delta = relaxApproximateInelasticCoulombContactsByDecoupling(dtinv); break;
      case ORC_INELASTIC_COULOMB_DECOUPLING:
        delta = relaxInelasticCoulombContactsByDecoupling(dtinv); break;
      case ORC_APPROX_INELASTIC_COULOMB_ORTHO:
        delta = relaxInelasticCoulombContactsByOrthogonalProjections(dtinv, true); break;
      case ORC_INELASTIC_COULOMB_ORTHO:
        delta = relaxInelasticCoulombContactsByOrthogonalProjections(dtinv, false); break;
      case ORC_INELASTIC_GMD:
        delta = relaxInelasticGeneralizedMaximumDissipationContacts(dtinv); break;
      default: throw std::runtime_error("Unsupported relaxation model.");
    }
    if (abortThreshold > 0 && delta < abortThreshold) { ++it; break; }
  }
  stats.iterations = it;
  stats.last_delta = delta;
  stats.t_solving = now() - t0;
  t0 = now();
  // Integration, HCSITS.h:1954-2004.  Bodies are local (no global section in a serial
  // run), so every body integrates with v + dv; for fixed bodies dv stays zero.
  for (size_t j = 0; j < nb; ++j) integratePositions(bodies[j], v_[j] + dv_[j], w_[j] + dw_[j], dt);
  stats.t_integration = now() - t0;
}

// HCSITS.h:2330-2595
double orc_world::relaxInelasticCoulombContactsByDecoupling(double dtinv) {
  double delta_max = 0;
  const size_t nc = p_.size();
  for (size_t i = 0; i < nc; ++i) {
    const int i1 = body1_[i], i2 = body2_[i];
    const Body& b1 = bodies[i1];
    const Body& b2 = bodies[i2];
    const M3 I1 = b1.invInertiaWF(), I2 = b2.invInertiaWF();
    dv_[i1] -= b1.invMass * p_[i];
    dw_[i1] -= I1 * cross(r1_[i], p_[i]);
    dv_[i2] += b2.invMass * p_[i];
    dw_[i2] += I2 * cross(r2_[i], p_[i]);

    const V3 gdot = (v_[i1] + dv_[i1]) - (v_[i2] + dv_[i2]) + cross(w_[i1] + dw_[i1], r1_[i]) -
                    cross(w_[i2] + dw_[i2], r2_[i]);
    const M3 cf = M3::cols(n_[i], t_[i], o_[i]);
    V3 gdot_nto = trans(cf) * gdot;
    gdot_nto[0] += dist_[i] * dtinv;

    if (gdot_nto[0] >= 0) {
      delta_max = std::max(delta_max, std::max(std::fabs(p_[i][0]), std::max(std::fabs(p_[i][1]), std::fabs(p_[i][2]))));
      p_[i] = V3();
    } else {
      V3 p_cf = -(diag_nto_inv_[i] * gdot_nto);
      double flimit = mu_[i] * p_cf[0];
      double fsq = p_cf[1] * p_cf[1] + p_cf[2] * p_cf[2];
      if (fsq > flimit * flimit || p_cf[0] < 0) {
        for (int j = 0; j < 20; ++j) {
          p_cf[0] = 0;
          V3 gc = gdot + cross(I1 * cross(r1_[i], t_[i] * p_cf[1] + o_[i] * p_cf[2]), r1_[i]) +
                  cross(I2 * cross(r2_[i], t_[i] * p_cf[1] + o_[i] * p_cf[2]), r2_[i]);
          const double gc_n = dot(n_[i], gc) + dist_[i] * dtinv;
          p_cf[0] = std::max(0.0, -(diag_n_inv_[i] * gc_n));
          p_cf[1] = p_cf[2] = 0.0;
          gc = gdot + cross(I1 * cross(r1_[i], n_[i] * p_cf[0]), r1_[i]) +
               cross(I2 * cross(r2_[i], n_[i] * p_cf[0]), r2_[i]);
          const double g_t = dot(t_[i], gc), g_o = dot(o_[i], gc);
          const M2& Dti = diag_to_inv_[i];
          // -(Dto^-1 * g), Matrix2x2.h:1904-1905
          const double r0 = -(Dti.m[0] * g_t + Dti.m[1] * g_o);
          const double r1 = -(Dti.m[2] * g_t + Dti.m[3] * g_o);
          p_cf[1] = r0;
          p_cf[2] = r1;
          flimit = mu_[i] * p_cf[0];
          fsq = p_cf[1] * p_cf[1] + p_cf[2] * p_cf[2];
          if (fsq > flimit * flimit) {
            const M3& D = diag_nto_[i];
            const double A00 = D[4], A01 = D[5], A10 = D[7], A11 = D[8];
            const double f = flimit / std::sqrt(fsq);
            const double phi = 0.5 * (1.0 + std::sqrt(5.0));
            const double shift = t_atan2(-p_cf[2], p_cf[1], crTrig);
            const double acos_f = t_acos(f, crTrig);
            double alpha_left = -acos_f - shift;
            double alpha_right = acos_f - shift;
            double alpha_mid = (alpha_right + alpha_left * phi) / (1 + phi);
            double xm0 = flimit * t_cos(alpha_mid, crTrig), xm1 = flimit * t_sin(alpha_mid, crTrig);
            // 0.5 * x^T (A x) - x^T (-g)
            double f_mid = 0.5 * xm0 * (A00 * xm0 + A01 * xm1) + 0.5 * xm1 * (A10 * xm0 + A11 * xm1) -
                           (xm0 * (-g_t) + xm1 * (-g_o));
            bool leftlarger = false;
            for (size_t k = 0; k < maxSubIterations; ++k) {
              const double alpha_next = alpha_left + (alpha_right - alpha_mid);
              const double xn0 = flimit * t_cos(alpha_next, crTrig), xn1 = flimit * t_sin(alpha_next, crTrig);
              const double f_next = 0.5 * xn0 * (A00 * xn0 + A01 * xn1) + 0.5 * xn1 * (A10 * xn0 + A11 * xn1) -
                                    (xn0 * (-g_t) + xn1 * (-g_o));
              if (leftlarger) {
                if (f_next < f_mid) {
                  alpha_right = alpha_mid; alpha_mid = alpha_next; xm0 = xn0; xm1 = xn1; f_mid = f_next;
                  leftlarger = true;
                } else {
                  alpha_left = alpha_next; leftlarger = false;
                }
              } else {
                if (f_next < f_mid) {
                  alpha_left = alpha_mid; alpha_mid = alpha_next; xm0 = xn0; xm1 = xn1; f_mid = f_next;
                  leftlarger = false;
                } else {
                  alpha_right = alpha_next; leftlarger = true;
                }
              }
            }
            p_cf[1] = xm0;
            p_cf[2] = xm1;
          }
        }
      }
      const V3 p_wf = cf * p_cf;
      const V3 dp = p_[i] - p_wf;
      delta_max = std::max(delta_max, std::max(std::fabs(dp[0]), std::max(std::fabs(dp[1]), std::fabs(dp[2]))));
      p_[i] = p_wf;
      dv_[i1] += b1.invMass * p_[i];
      dw_[i1] += I1 * cross(r1_[i], p_[i]);
      dv_[i2] -= b2.invMass * p_[i];
      dw_[i2] -= I2 * cross(r2_[i], p_[i]);
    }
  }
  return delta_max;
}

// HCSITS.h:2129-2304
double orc_world::relaxApproximateInelasticCoulombContactsByDecoupling(double dtinv) {
  double delta_max = 0;
  const size_t nc = p_.size();
  for (size_t i = 0; i < nc; ++i) {
    const int i1 = body1_[i], i2 = body2_[i];
    const Body& b1 = bodies[i1];
    const Body& b2 = bodies[i2];
    const M3 I1 = b1.invInertiaWF(), I2 = b2.invInertiaWF();
    dv_[i1] -= b1.invMass * p_[i];
    dw_[i1] -= I1 * cross(r1_[i], p_[i]);
    dv_[i2] += b2.invMass * p_[i];
    dw_[i2] += I2 * cross(r2_[i], p_[i]);
    const V3 gdot = (v_[i1] + dv_[i1]) - (v_[i2] + dv_[i2]) + cross(w_[i1] + dw_[i1], r1_[i]) -
                    cross(w_[i2] + dw_[i2], r2_[i]);
    const M3 cf = M3::cols(n_[i], t_[i], o_[i]);
    V3 gdot_nto = trans(cf) * gdot;
    gdot_nto[0] += dist_[i] * dtinv;
    if (gdot_nto[0] >= 0) {
      delta_max = std::max(delta_max, std::max(std::fabs(p_[i][0]), std::max(std::fabs(p_[i][1]), std::fabs(p_[i][2]))));
      p_[i] = V3();
    } else {
      V3 p_cf = -(diag_nto_inv_[i] * gdot_nto);
      const double flimit = mu_[i] * p_cf[0];
      const double fsq = p_cf[1] * p_cf[1] + p_cf[2] * p_cf[2];
      if (fsq > flimit * flimit || p_cf[0] < 0) {
        V3 p_tmp = dot(t_[i], p_[i]) * t_[i] + dot(o_[i], p_[i]) * o_[i];
        const double gdot_n = gdot_nto[0] + dot(n_[i], cross(I1 * cross(r1_[i], p_tmp), r1_[i]) +
                                                           cross(I2 * cross(r2_[i], p_tmp), r2_[i]));
        p_cf[0] = -(diag_n_inv_[i] * gdot_n);
        p_cf[0] = std::max(0.0, p_cf[0]);
        p_tmp = n_[i] * p_cf[0];
        const V3 gc = gdot + cross(I1 * cross(r1_[i], p_tmp), r1_[i]) + cross(I2 * cross(r2_[i], p_tmp), r2_[i]);
        const double g_t = dot(t_[i], gc), g_o = dot(o_[i], gc);
        const M2& Dti = diag_to_inv_[i];
        p_cf[1] = -(Dti.m[0] * g_t + Dti.m[1] * g_o);
        p_cf[2] = -(Dti.m[2] * g_t + Dti.m[3] * g_o);
        const double fl = mu_[i] * p_cf[0];
        const double fs = p_cf[1] * p_cf[1] + p_cf[2] * p_cf[2];
        if (fs > fl * fl) {
          const double f = fl / std::sqrt(fs);
          p_cf[1] *= f;
          p_cf[2] *= f;
        }
      }
      const V3 p_wf = cf * p_cf;
      const V3 dp = p_[i] - p_wf;
      delta_max = std::max(delta_max, std::max(std::fabs(dp[0]), std::max(std::fabs(dp[1]), std::fabs(dp[2]))));
      p_[i] = p_wf;
      dv_[i1] += b1.invMass * p_[i];
      dw_[i1] += I1 * cross(r1_[i], p_[i]);
      dv_[i2] -= b2.invMass * p_[i];
      dw_[i2] -= I2 * cross(r2_[i], p_[i]);
    }
  }
  return delta_max;
}

// HCSITS.h:2631-2723: the impulse in the contact frame takes one step p - (D p + gdot_nto)
// (approximate: Anitescu/Tasora) or p - (D p + gdot_nto + mu |(D p + gdot_nto)_to| e_n)
// (De Saxce/Feng) and is projected onto the friction cone by the reference's three cases.
double orc_world::relaxInelasticCoulombContactsByOrthogonalProjections(double dtinv, bool approximate) {
  double delta_max = 0;
  const size_t nc = p_.size();
  for (size_t i = 0; i < nc; ++i) {
    const int i1 = body1_[i], i2 = body2_[i];
    const Body& b1 = bodies[i1];
    const Body& b2 = bodies[i2];
    const M3 I1 = b1.invInertiaWF(), I2 = b2.invInertiaWF();
    dv_[i1] -= b1.invMass * p_[i];
    dw_[i1] -= I1 * cross(r1_[i], p_[i]);
    dv_[i2] += b2.invMass * p_[i];
    dw_[i2] += I2 * cross(r2_[i], p_[i]);
    const V3 gdot = (v_[i1] + dv_[i1]) - (v_[i2] + dv_[i2]) + cross(w_[i1] + dw_[i1], r1_[i]) -
                    cross(w_[i2] + dw_[i2], r2_[i]);
    const M3 cf = M3::cols(n_[i], t_[i], o_[i]);
    V3 gdot_nto = trans(cf) * gdot;
    gdot_nto[0] += dist_[i] * dtinv;
    const double step = 1.0;   // the reference's w
    const double mu = mu_[i];
    V3 p_cf = trans(cf) * p_[i];
    V3 corr = diag_nto_[i] * p_cf + gdot_nto;
    if (!approximate) corr[0] += std::sqrt(corr[1] * corr[1] + corr[2] * corr[2]) * mu;
    p_cf = p_cf - step * corr;
    const double flimit = mu * p_cf[0];
    const double fsq = p_cf[1] * p_cf[1] + p_cf[2] * p_cf[2];
    if (p_cf[0] > 0 && fsq < flimit * flimit) {
      // static: no projection
    } else if (p_cf[0] < 0 && fsq < p_cf[0] / (mu * mu)) {
      p_cf = V3();
    } else {
      const double fn = std::sqrt(fsq);
      p_cf[0] = (fn * mu + p_cf[0]) / (mu * mu + 1);
      const double factor = mu * p_cf[0] / fn;
      p_cf[1] *= factor;
      p_cf[2] *= factor;
    }
    const V3 p_wf = cf * p_cf;
    const V3 dp = p_[i] - p_wf;
    delta_max = std::max(delta_max, std::max(std::fabs(dp[0]), std::max(std::fabs(dp[1]), std::fabs(dp[2]))));
    p_[i] = p_wf;
    dv_[i1] += b1.invMass * p_[i];
    dw_[i1] += I1 * cross(r1_[i], p_[i]);
    dv_[i2] -= b2.invMass * p_[i];
    dw_[i2] -= I2 * cross(r2_[i], p_[i]);
  }
  return delta_max;
}

// HCSITS.h:2745-3055 (relaxInelasticGeneralizedMaximumDissipationContacts).  A sliding contact
// takes the impulse on the friction cone's boundary that satisfies the normal constraint
// (gdot_n + D p = 0 in the normal row) and minimises the energy 1/2 x^T A_ x - x^T b_ of the
// tangential part x, with A_ and b_ the tangential block and load after eliminating the normal
// row.  On the cone boundary |x| = mu p_n, so x = r(alpha) (cos alpha, sin alpha) with
// r(alpha) = mu (-gdot_n) / (D00 + mu a3 cos(alpha + shiftI)), and a golden-section search over
// alpha finds the minimum; the bracket is where r(alpha) is positive (angleI) intersected with
// where the cone boundary is cheaper than the unconstrained direction (angleJ).
double orc_world::relaxInelasticGeneralizedMaximumDissipationContacts(double dtinv) {
  const double kPi = 3.141592653589793;   // M_PI (pe/math/Constants.h:163)
  double delta_max = 0;
  const size_t nc = p_.size();
  for (size_t i = 0; i < nc; ++i) {
    const int i1 = body1_[i], i2 = body2_[i];
    const Body& b1 = bodies[i1];
    const Body& b2 = bodies[i2];
    const M3 I1 = b1.invInertiaWF(), I2 = b2.invInertiaWF();
    dv_[i1] -= b1.invMass * p_[i];
    dw_[i1] -= I1 * cross(r1_[i], p_[i]);
    dv_[i2] += b2.invMass * p_[i];
    dw_[i2] += I2 * cross(r2_[i], p_[i]);
    const V3 gdot = (v_[i1] + dv_[i1]) - (v_[i2] + dv_[i2]) + cross(w_[i1] + dw_[i1], r1_[i]) -
                    cross(w_[i2] + dw_[i2], r2_[i]);
    const M3 cf = M3::cols(n_[i], t_[i], o_[i]);
    V3 gdot_nto = trans(cf) * gdot;
    gdot_nto[0] += dist_[i] * dtinv;
    if (gdot_nto[0] >= 0) {
      delta_max = std::max(delta_max, std::max(std::fabs(p_[i][0]), std::max(std::fabs(p_[i][1]), std::fabs(p_[i][2]))));
      p_[i] = V3();
      continue;
    }
    V3 p_cf = -(diag_nto_inv_[i] * gdot_nto);
    const double mu = mu_[i];
    const double flimit = mu * p_cf[0];
    const double fsq = p_cf[1] * p_cf[1] + p_cf[2] * p_cf[2];
    if (fsq > flimit * flimit || p_cf[0] < 0) {
      const M3& D = diag_nto_[i];
      const double D00 = D[0], D01 = D[1], D02 = D[2], D11 = D[4], D12 = D[5], D22 = D[8];
      const double invA_nn = 1.0 / D00;   // inv(real), pe/math/shims/Invert.h:62-65
      const double offdiag = D12 - invA_nn * D01 * D02;
      // Mat2 A_breve (row-major) and Vec2 b_breve (HCSITS.h:2831-2832)
      const double A00 = D11 - invA_nn * (D01 * D01), A01 = offdiag, A10 = offdiag,
                   A11 = D22 - invA_nn * (D02 * D02);
      const double bb0 = -gdot_nto[1] + invA_nn * D01 * gdot_nto[0];
      const double bb1 = -gdot_nto[2] + invA_nn * D02 * gdot_nto[0];
      const double shiftI = t_atan2(-D02, D01, crTrig);
      const double shiftJ = t_atan2(-p_cf[2], p_cf[1], crTrig);
      const double a3 = std::sqrt(D01 * D01 + D02 * D02);
      const double fractionI = -D00 / (mu * a3);
      const double fj = invA_nn * mu * ((-gdot_nto[0]) / std::sqrt(fsq) - a3 * t_cos(shiftI - shiftJ, crTrig));
      const double fractionJ = std::min(fj, 1.0);
      double alpha_left, alpha_right;
      if (fractionJ < -1) {
        const double angleI = t_acos(fractionI, crTrig);
        alpha_left = -angleI - shiftI;
        alpha_right = +angleI - shiftI;
        if (alpha_left < 0) { alpha_left += 2 * kPi; alpha_right += 2 * kPi; }
      } else if (D00 > mu * a3) {
        const double angleJ = t_acos(fractionJ, crTrig);
        alpha_left = -angleJ - shiftJ;
        alpha_right = +angleJ - shiftJ;
        if (alpha_left < 0) { alpha_left += 2 * kPi; alpha_right += 2 * kPi; }
      } else {
        const double angleJ = t_acos(fractionJ, crTrig);
        double a1l = -angleJ - shiftJ, a1r = +angleJ - shiftJ;
        if (a1l < 0) { a1l += 2 * kPi; a1r += 2 * kPi; }
        const double angleI = t_acos(fractionI, crTrig);
        double a2l = -angleI - shiftI, a2r = +angleI - shiftI;
        if (a2l < 0) { a2l += 2 * kPi; a2r += 2 * kPi; }
        if (a1l > a2l) { std::swap(a1l, a2l); std::swap(a1r, a2r); }
        if (a2l > a1r) {
          a2r -= 2 * kPi;
          if (!(a2r > a1r)) a1r = a2r;
        } else {
          a1l = a2l;
          if (!(a2r > a1r)) a1r = a2r;
        }
        alpha_left = a1l;
        alpha_right = a1r;
      }
      // a point of the cone boundary at angle alpha and its energy (HCSITS.h:2924-2932)
      auto eval = [&](double alpha, double& x0, double& x1) {
        double r_ub = mu * (-gdot_nto[0]) / (D00 + mu * a3 * t_cos(alpha + shiftI, crTrig));
        if (r_ub < 0) r_ub = kInf;
        x0 = r_ub * t_cos(alpha, crTrig);
        x1 = r_ub * t_sin(alpha, crTrig);
        return 0.5 * x0 * (A00 * x0 + A01 * x1) + 0.5 * x1 * (A10 * x0 + A11 * x1) - (x0 * bb0 + x1 * bb1);
      };
      const double phi = 0.5 * (1.0 + std::sqrt(5.0));
      double alpha_mid = (alpha_right + alpha_left * phi) / (1 + phi);
      double xm0, xm1;
      double f_mid = eval(alpha_mid, xm0, xm1);
      bool leftlarger = false;
      for (size_t k = 0; k < maxSubIterations; ++k) {
        const double alpha_next = alpha_left + (alpha_right - alpha_mid);
        double xn0, xn1;
        const double f_next = eval(alpha_next, xn0, xn1);
        if (leftlarger) {
          if (f_next < f_mid) {
            alpha_right = alpha_mid; alpha_mid = alpha_next; xm0 = xn0; xm1 = xn1; f_mid = f_next;
            leftlarger = true;
          } else {
            alpha_left = alpha_next; leftlarger = false;
          }
        } else {
          if (f_next < f_mid) {
            alpha_left = alpha_mid; alpha_mid = alpha_next; xm0 = xn0; xm1 = xn1; f_mid = f_next;
            leftlarger = false;
          } else {
            alpha_right = alpha_next; leftlarger = true;
          }
        }
      }
      // the direction of the unconstrained impulse, if cheaper (HCSITS.h:3001-3015)
      {
        const double alpha_init = t_atan2(p_cf[2], p_cf[1], crTrig);
        double xi0, xi1;
        const double f_init = eval(alpha_init, xi0, xi1);
        if (f_init < f_mid) { xm0 = xi0; xm1 = xi1; }
      }
      p_cf[0] = invA_nn * (-gdot_nto[0] - D01 * xm0 - D02 * xm1);
      p_cf[1] = xm0;
      p_cf[2] = xm1;
    }
    const V3 p_wf = cf * p_cf;
    const V3 dp = p_[i] - p_wf;
    delta_max = std::max(delta_max, std::max(std::fabs(dp[0]), std::max(std::fabs(dp[1]), std::fabs(dp[2]))));
    p_[i] = p_wf;
    dv_[i1] += b1.invMass * p_[i];
    dw_[i1] += I1 * cross(r1_[i], p_[i]);
    dv_[i2] -= b2.invMass * p_[i];
    dw_[i2] -= I2 * cross(r2_[i], p_[i]);
  }
  return delta_max;
}

// HCSITS.h:2039-2105
double orc_world::relaxInelasticFrictionlessContacts(double dtinv) {
  double delta_max = 0;
  const size_t nc = p_.size();
  for (size_t i = 0; i < nc; ++i) {
    const int i1 = body1_[i], i2 = body2_[i];
    const Body& b1 = bodies[i1];
    const Body& b2 = bodies[i2];
    const M3 I1 = b1.invInertiaWF(), I2 = b2.invInertiaWF();
    dv_[i1] -= b1.invMass * p_[i];
    dw_[i1] -= I1 * cross(r1_[i], p_[i]);
    dv_[i2] += b2.invMass * p_[i];
    dw_[i2] += I2 * cross(r2_[i], p_[i]);
    const V3 gdot = (v_[i1] + dv_[i1]) - (v_[i2] + dv_[i2]) + cross(w_[i1] + dw_[i1], r1_[i]) -
                    cross(w_[i2] + dw_[i2], r2_[i]);
    const M3 cf = M3::cols(n_[i], t_[i], o_[i]);
    V3 gdot_nto = trans(cf) * gdot;
    gdot_nto[0] += dist_[i] * dtinv;
    if (gdot_nto[0] >= 0) {
      delta_max = std::max(delta_max, std::max(std::fabs(p_[i][0]), std::max(std::fabs(p_[i][1]), std::fabs(p_[i][2]))));
      p_[i] = V3();
    } else {
      const V3 p_wf = n_[i] * (-diag_n_inv_[i] * gdot_nto[0]);
      const V3 dp = p_[i] - p_wf;
      delta_max = std::max(delta_max, std::max(std::fabs(dp[0]), std::max(std::fabs(dp[1]), std::fabs(dp[2]))));
      p_[i] = p_wf;
      dv_[i1] += b1.invMass * p_[i];
      dw_[i1] += I1 * cross(r1_[i], p_[i]);
      dv_[i2] -= b2.invMass * p_[i];
      dw_[i2] -= I2 * cross(r2_[i], p_[i]);
    }
  }
  return delta_max;
}

// =============================================================================================
// C API
// =============================================================================================
extern "C" {

orc_world* orc_create(void) { return new orc_world(); }
void orc_destroy(orc_world* w) { delete w; }
void orc_set_gravity(orc_world* w, double gx, double gy, double gz) { w->gravity = V3(gx, gy, gz); }
void orc_set_damping(orc_world* w, double damping) { w->damping = damping; }
void orc_set_max_iterations(orc_world* w, size_t n) { w->maxIterations = n; }
void orc_set_relaxation_model(orc_world* w, int model) { w->relaxationModel = model; }
void orc_set_erp(orc_world* w, double erp) { w->erp = erp; }
void orc_set_abort_threshold(orc_world* w, double thr) { w->abortThreshold = thr; }
void orc_set_trig_mode(orc_world* w, int correctly_rounded) { w->crTrig = correctly_rounded != 0; }
double orc_trig(int fn, double x, double y, int correctly_rounded) {
  const bool cr = correctly_rounded != 0;
  switch (fn) {
    case 0: return t_sin(x, cr);
    case 1: return t_cos(x, cr);
    case 2: return t_atan2(y, x, cr);
    default: return t_acos(x, cr);
  }
}

// createMaterial, Materials.cpp:175-247 (pair tables: csf(i,j) = csf_i + csf_j)
int orc_create_material(orc_world* w, double density, double cor, double csf, double cdf) {
  Material m{density, cor, csf, cdf};
  const size_t n0 = w->materials.size();
  w->materials.push_back(m);
  const size_t n = n0 + 1;
  std::vector<double> ns(n * n, 0.0), nd(n * n, 0.0);
  for (size_t i = 0; i < n0; ++i)
    for (size_t j = 0; j < n0; ++j) {
      ns[i * n + j] = w->csf[i * n0 + j];
      nd[i * n + j] = w->cdf[i * n0 + j];
    }
  ns[n0 * n + n0] = csf + csf;
  nd[n0 * n + n0] = cdf + cdf;
  for (size_t i = 0; i < n0; ++i) {
    ns[n0 * n + i] = ns[i * n + n0] = w->materials[i].csf + csf;
    nd[n0 * n + i] = nd[i * n + n0] = w->materials[i].cdf + cdf;
  }
  w->csf.swap(ns);
  w->cdf.swap(nd);
  return (int)n0;
}

void orc_set_pair_friction(orc_world* w, int m1, int m2, double csf, double cdf) {
  const size_t n = w->materials.size();
  w->csf[(size_t)m1 * n + m2] = w->csf[(size_t)m2 * n + m1] = csf;
  w->cdf[(size_t)m1 * n + m2] = w->cdf[(size_t)m2 * n + m1] = cdf;
}

static int register_body(orc_world* w, Body& b) {
  b.sid = w->nextSid++;  // UniqueID<RigidBody>::create, UniqueID.h:143
  w->bodies.push_back(b);
  const int idx = (int)w->bodies.size() - 1;
  w->hgAdd(idx);  // World::add -> CollisionSystem::add -> HashGrids::add (HCSITS.h:1305-1325)
  return idx;
}

// createSphere / SphereBase ctor, SphereBase.cpp:64-87, SphereBase.h:169-234
int orc_create_sphere(orc_world* w, double x, double y, double z, double r, int mat) {
  Body b;
  b.shape = ORC_SPHERE;
  b.mat = mat;
  b.radius = r;
  b.gpos = V3(x, y, z);
  b.mass = 4.0 / 3.0 * M_PI * r * r * r * w->materials[mat].density;
  b.invMass = 1.0 / b.mass;
  b.I = M3();
  b.I[0] = b.I[4] = b.I[8] = 0.4 * b.mass * r * r;
  b.Iinv = inverse(b.I);
  b.calcBoundingBox();
  return register_body(w, b);
}

// createBox / BoxBase ctor, BoxBase.cpp:60-160, BoxBase.h:186-203
int orc_create_box(orc_world* w, double x, double y, double z, double lx, double ly, double lz, int mat) {
  Body b;
  b.shape = ORC_BOX;
  b.mat = mat;
  b.lengths = V3(lx, ly, lz);
  b.gpos = V3(x, y, z);
  b.mass = lx * ly * lz * w->materials[mat].density;
  b.invMass = 1.0 / b.mass;
  b.I = M3();
  b.I[0] = b.mass / 12.0 * (ly * ly + lz * lz);
  b.I[4] = b.mass / 12.0 * (lx * lx + lz * lz);
  b.I[8] = b.mass / 12.0 * (lx * lx + ly * ly);
  b.Iinv = inverse(b.I);
  b.calcBoundingBox();
  return register_body(w, b);
}

// createCapsule / CapsuleBase ctor, CapsuleBase.cpp:60-92, 147-168; CapsuleBase.h:182-200.  The
// axis is the body x-axis; mass = pi r^2 (4/3 r + length) rho; the moments combine the cylinder
// and the two half-sphere caps, with the density recomputed from the mass as the reference does.
int orc_create_capsule(orc_world* w, double x, double y, double z, double r, double len, int mat) {
  Body b;
  b.shape = ORC_CAPSULE;
  b.mat = mat;
  b.radius = r;
  b.length = len;
  b.gpos = V3(x, y, z);
  b.mass = r * r * M_PI * (4.0 / 3.0 * r + len) * w->materials[mat].density;
  b.invMass = 1.0 / b.mass;
  const double density = b.mass / (M_PI * r * r * (4.0 / 3.0 * r + len));
  const double sphereMass = 4.0 / 3.0 * M_PI * r * r * r * density;
  const double cylinderMass = M_PI * r * r * len * density;
  const double Ia = r * r * (0.5 * cylinderMass + 0.4 * sphereMass);
  const double Ib = cylinderMass * (0.25 * r * r + 1.0 / 12.0 * len * len) +
                    sphereMass * (0.4 * r * r + 0.375 * r * len + 0.25 * len * len);
  b.I = M3();
  b.I[0] = Ia;
  b.I[4] = Ib;
  b.I[8] = Ib;
  b.Iinv = inverse(b.I);
  b.calcBoundingBox();
  return register_body(w, b);
}

// A union of spheres as pe builds one inside pe_CREATE_UNION and places it with setPosition
// (createGranularParticle, Union.cpp:2376-2402): the union takes the next system ID, then each
// sphere is created at its offset disp[k] (its own system ID; SphereBase's mass properties) and
// added (Union::add, Union.cpp:1482-1545).  After the last add: the total mass and centre of mass
// (UnionBase::calcCenterOfMass, UnionBase.cpp:240-313), the members' relative positions
// (RigidBody::calcRelPosition, RigidBody.h:1370-1373) and the moment of inertia about the centre of
// mass, rotated into the body frame (UnionBase::calcInertia, UnionBase.cpp:323-360), all with the
// union's identity orientation.  Union::setPosition (Union.cpp:251-272) then moves the centre of
// mass to (x, y, z) and every member by the same displacement.
int orc_create_union(orc_world* w, double x, double y, double z, int n, const double* disp,
                     const double* radius, const int* mat) {
  if (n < 1) return -1;
  Body u;
  u.shape = ORC_UNION;
  u.mat = mat[0];
  u.sid = w->nextSid++;
  u.mbegin = (int)w->members.size();
  u.mcount = n;
  const int uidx = (int)w->bodies.size();
  for (int k = 0; k < n; ++k) {
    Body m;
    m.shape = ORC_SPHERE;
    m.mat = mat[k];
    m.radius = radius[k];
    m.gpos = V3(disp[3 * k], disp[3 * k + 1], disp[3 * k + 2]);
    m.mass = 4.0 / 3.0 * M_PI * m.radius * m.radius * m.radius * w->materials[mat[k]].density;
    m.invMass = 1.0 / m.mass;
    m.I = M3();
    m.I[0] = m.I[4] = m.I[8] = 0.4 * m.mass * m.radius * m.radius;
    m.Iinv = inverse(m.I);
    m.sid = w->nextSid++;
    m.super = uidx;
    w->members.push_back(m);
  }
  Body* ms = &w->members[u.mbegin];
  if (n == 1) {
    u.gpos = ms[0].gpos;
    u.mass = ms[0].mass;
    u.invMass = 1.0 / u.mass;
  } else {
    u.gpos = V3();
    u.mass = 0;
    for (int k = 0; k < n; ++k) {
      u.mass += ms[k].mass;
      u.gpos += ms[k].gpos * ms[k].mass;
    }
    u.invMass = 1.0 / u.mass;
    const double tmp = 1.0 / u.mass;  // Vector3::operator/=, Vector3.h:836-848
    u.gpos = V3(u.gpos[0] * tmp, u.gpos[1] * tmp, u.gpos[2] * tmp);
  }
  for (int k = 0; k < n; ++k) ms[k].rpos = trans(u.R) * (ms[k].gpos - u.gpos);
  M3 I;  // zero
  for (int k = 0; k < n; ++k) {
    const double mass = ms[k].mass;
    const V3 pos = ms[k].gpos - u.gpos;
    for (int j = 0; j < 9; ++j) I[j] += ms[k].I[j];
    I[0] += mass * (pos[1] * pos[1] + pos[2] * pos[2]);
    I[1] -= mass * pos[0] * pos[1];
    I[2] -= mass * pos[0] * pos[2];
    I[3] -= mass * pos[0] * pos[1];
    I[4] += mass * (pos[0] * pos[0] + pos[2] * pos[2]);
    I[5] -= mass * pos[1] * pos[2];
    I[6] -= mass * pos[0] * pos[2];
    I[7] -= mass * pos[1] * pos[2];
    I[8] += mass * (pos[0] * pos[0] + pos[1] * pos[1]);
  }
  u.I = trans(u.R) * I * u.R;
  u.Iinv = inverse(u.I);
  const V3 gpos(x, y, z);
  const V3 dp = gpos - u.gpos;
  u.gpos = gpos;
  for (int k = 0; k < n; ++k) {
    ms[k].gpos += dp;  // Sphere::update(dp), Sphere.cpp:1024-1038
    ms[k].calcBoundingBox();
  }
  w->calcBoundingBox(u);
  w->bodies.push_back(u);
  w->hgAdd(uidx);
  return uidx;
}

size_t orc_num_members(const orc_world* w) { return w->members.size(); }

void orc_get_members(const orc_world* w, uint32_t* owner, uint64_t* sid, int* mat, double* radius,
                     double* rpos, double* gpos) {
  for (size_t k = 0; k < w->members.size(); ++k) {
    const Body& m = w->members[k];
    if (owner) owner[k] = (uint32_t)m.super;
    if (sid) sid[k] = m.sid;
    if (mat) mat[k] = m.mat;
    if (radius) radius[k] = m.radius;
    for (int j = 0; j < 3; ++j) {
      if (rpos) rpos[3 * k + j] = m.rpos[j];
      if (gpos) gpos[3 * k + j] = m.gpos[j];
    }
  }
}

void orc_set_members(orc_world* w, const double* gpos) {
  for (size_t k = 0; k < w->members.size(); ++k) {
    Body& m = w->members[k];
    m.gpos = V3(gpos[3 * k], gpos[3 * k + 1], gpos[3 * k + 2]);
    m.calcBoundingBox();
  }
  for (Body& b : w->bodies)
    if (b.shape == ORC_UNION) w->calcBoundingBox(b);
}

// createPlane, Plane.cpp:1202-1240; Plane ctor, Plane.cpp:77-102; PlaneBase.cpp:73-100
int orc_create_plane(orc_world* w, double nx, double ny, double nz, double d, int mat) {
  Body b;
  b.shape = ORC_PLANE;
  b.mat = mat;
  V3 n = normalized(V3(nx, ny, nz));
  b.normal = n;
  b.d = d;
  b.gpos = n * d;
  b.fixed = true;
  b.mass = 0; b.invMass = 0;
  if (n[0] * n[0] + n[1] * n[1] < kAccuracy)
    b.q = quat_axis_angle(V3(1, 0, 0), std::acos(n[2]));
  else
    b.q = quat_axis_angle(V3(-n[1], n[0], 0.0), std::acos(n[2]));
  b.R = quat_to_R(b.q);
  b.calcBoundingBox();
  return register_body(w, b);
}

size_t orc_num_bodies(const orc_world* w) { return w->bodies.size(); }

void orc_get_state(const orc_world* w, double* pos, double* q, double* v, double* om,
                   double* normal, double* d) {
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    const Body& b = w->bodies[i];
    for (int k = 0; k < 3; ++k) {
      pos[3 * i + k] = b.gpos[k];
      v[3 * i + k] = b.v[k];
      om[3 * i + k] = b.w[k];
      normal[3 * i + k] = b.normal[k];
    }
    for (int k = 0; k < 4; ++k) q[4 * i + k] = b.q.v[k];
    d[i] = b.d;
  }
}

void orc_set_state(orc_world* w, const double* pos, const double* q, const double* v,
                   const double* om, const double* normal, const double* d) {
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    Body& b = w->bodies[i];
    for (int k = 0; k < 3; ++k) {
      b.gpos[k] = pos[3 * i + k];
      b.v[k] = v[3 * i + k];
      b.w[k] = om[3 * i + k];
    }
    for (int k = 0; k < 4; ++k) b.q.v[k] = q[4 * i + k];
    b.R = quat_to_R(b.q);
    if (b.shape == ORC_PLANE) {
      for (int k = 0; k < 3; ++k) b.normal[k] = normal[3 * i + k];
      b.d = d[i];
    }
    w->calcBoundingBox(b);  // a union's members keep their positions (orc_set_members moves them)
  }
}

// RigidBody::addForce / addTorque (RigidBody.h:262-274): applied at the next step's body caching, then
// reset (Settings::forceReset, Settings.cpp:46).
void orc_set_forces(orc_world* w, const double* force, const double* torque) {
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    Body& b = w->bodies[i];
    for (int k = 0; k < 3; ++k) {
      b.force[k] = force[3 * i + k];
      b.torque[k] = torque[3 * i + k];
    }
  }
}

void orc_get_body_info(const orc_world* w, int* shape, uint64_t* sid, int* mat, int* fixed,
                       double* param, double* inv_mass, double* inv_inertia_body) {
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    const Body& b = w->bodies[i];
    shape[i] = b.shape;
    sid[i] = b.sid;
    mat[i] = b.mat;
    fixed[i] = b.fixed ? 1 : 0;
    if (b.shape == ORC_SPHERE) { param[3 * i] = b.radius; param[3 * i + 1] = 0; param[3 * i + 2] = 0; }
    else if (b.shape == ORC_BOX) { for (int k = 0; k < 3; ++k) param[3 * i + k] = b.lengths[k]; }
    else if (b.shape == ORC_CAPSULE) { param[3 * i] = b.radius; param[3 * i + 1] = b.length; param[3 * i + 2] = 0; }
    else if (b.shape == ORC_UNION) { param[3 * i] = b.mbegin; param[3 * i + 1] = b.mcount; param[3 * i + 2] = 0; }
    else { param[3 * i] = param[3 * i + 1] = param[3 * i + 2] = 0; }
    inv_mass[i] = b.invMass;
    for (int k = 0; k < 9; ++k) inv_inertia_body[9 * i + k] = b.Iinv[k];
  }
}

void orc_get_aabb(const orc_world* w, double* aabb) {
  for (size_t i = 0; i < w->bodies.size(); ++i)
    for (int k = 0; k < 6; ++k) aabb[6 * i + k] = w->bodies[i].aabb[k];
}

int orc_step(orc_world* w, double dt, int nsteps) {
  if (dt <= 0) return -1;  // World::simulationStep, World.cpp:333-334
  try {
    for (int s = 0; s < nsteps; ++s) w->simulationStep(dt);
  } catch (const std::exception&) {
    return -2;
  }
  return 0;
}

void orc_set_sweep_order(orc_world* w, size_t n, const uint32_t* b1, const uint32_t* b2) {
  w->orderB1.assign(b1, b1 + n);
  w->orderB2.assign(b2, b2 + n);
  w->orderRank.clear();
  w->orderPending = true;
}

void orc_set_sweep_order_ranked(orc_world* w, size_t n, const uint32_t* b1, const uint32_t* b2,
                                const uint32_t* rank) {
  orc_set_sweep_order(w, n, b1, b2);
  w->orderRank.assign(rank, rank + n);
}

size_t orc_num_contacts(const orc_world* w) { return w->contacts.size(); }

void orc_get_contacts(const orc_world* w, uint32_t* b1, uint32_t* b2, double* pos,
                      double* normal, double* dist, double* mu) {
  for (size_t i = 0; i < w->contacts.size(); ++i) {
    const Contact& c = w->contacts[i];
    b1[i] = (uint32_t)c.b1;
    b2[i] = (uint32_t)c.b2;
    for (int k = 0; k < 3; ++k) { pos[3 * i + k] = c.pos[k]; normal[3 * i + k] = c.normal[k]; }
    dist[i] = c.dist;
    mu[i] = c.mu;
  }
}

void orc_get_impulses(const orc_world* w, double* p) {
  for (size_t i = 0; i < w->p_.size(); ++i)
    for (int k = 0; k < 3; ++k) p[3 * i + k] = w->p_[i][k];
}

size_t orc_detect(orc_world* w, int exhaustive) {
  w->contacts.clear();
  w->p_.clear();   // no solve has run on the contacts found here: their impulses are zero
  Fine fine{w, &w->contacts};
  try {
    if (exhaustive) w->exhaustiveFindContacts(fine);
    else w->hgFindContacts(fine);
  } catch (const std::exception&) {
    return (size_t)-1;
  }
  w->p_.assign(w->contacts.size(), V3{});
  return w->contacts.size();
}

size_t orc_candidate_pairs(const orc_world* w, uint32_t* a, uint32_t* b, size_t cap) {
  size_t n = 0;
  for (size_t i = 0; i < w->bodies.size(); ++i)
    for (size_t j = i + 1; j < w->bodies.size(); ++j) {
      const Body& A = w->bodies[i];
      const Body& B = w->bodies[j];
      if ((!A.fixed || !B.fixed) && overlaps(A.aabb, B.aabb)) {
        if (n < cap) { a[n] = (uint32_t)i; b[n] = (uint32_t)j; }
        ++n;
      }
    }
  return n;
}

void orc_get_stats(const orc_world* w, orc_stats* s) { *s = w->stats; }

void orc_quat_from_axis_angle(const double axis[3], double angle, double q[4]) {
  const Q r = quat_axis_angle(V3(axis[0], axis[1], axis[2]), angle);
  for (int k = 0; k < 4; ++k) q[k] = r.v[k];
}
void orc_quat_to_rotation(const double q[4], double R[9]) {
  Q a; for (int k = 0; k < 4; ++k) a.v[k] = q[k];
  const M3 m = quat_to_R(a);
  for (int k = 0; k < 9; ++k) R[k] = m[k];
}
void orc_quat_mul(const double a[4], const double b[4], double out[4]) {
  Q x, y; for (int k = 0; k < 4; ++k) { x.v[k] = a[k]; y.v[k] = b[k]; }
  const Q r = quat_mul(x, y);
  for (int k = 0; k < 4; ++k) out[k] = r.v[k];
}
void orc_quat_rotate(const double q[4], const double v[3], double out[3]) {
  Q a; for (int k = 0; k < 4; ++k) a.v[k] = q[k];
  const V3 r = quat_rotate(a, V3(v[0], v[1], v[2]));
  for (int k = 0; k < 3; ++k) out[k] = r[k];
}
void orc_mat3_inverse(const double m[9], double out[9]) {
  M3 a; for (int k = 0; k < 9; ++k) a[k] = m[k];
  const M3 r = inverse(a);
  for (int k = 0; k < 9; ++k) out[k] = r[k];
}

}  // extern "C"
