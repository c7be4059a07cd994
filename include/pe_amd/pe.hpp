// pe_amd/pe.hpp -- C++ drop-in for pe's World / CollisionSystem API on the HIP time step.
//
// A pe program that uses the hard-contact solver (pe_CONSTRAINT_SOLVER = HCSITS) switches
// by replacing `#include <pe/engine.h>` with `#include <pe_amd/pe.hpp>` and linking
// libpehip.so.  Names, argument order and error behaviour follow the reference:
//   theWorld(), World::run/simulationStep/setGravity/setDamping   (pe/core/World.h, src/core/World.cpp:330-378)
//   theCollisionSystem() and its HCSITS setters/getters            (HCSITS.h:198-260, 715-840)
//   createMaterial                                                  (src/core/Materials.cpp:175-247)
//   createSphere / createBox / createCapsule / createPlane          (Sphere.cpp:1187, Box.cpp:1312,
//                                                                    Capsule.cpp:1219, Plane.cpp:1202)
//   createUnion / createGranularParticle                            (Union.cpp:1960, 2376)
// Body handles (SphereID, BoxID, PlaneID) read and write their state through the device;
// the state is downloaded lazily after a step and uploaded before the next one.
#ifndef PE_AMD_PE_HPP
#define PE_AMD_PE_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../pe_hip.h"

namespace pe {

typedef double real;
typedef uint64_t id_t;

struct Vec3 {
  real v[3];
  Vec3() : v{0, 0, 0} {}
  Vec3(real x, real y, real z) : v{x, y, z} {}
  real& operator[](size_t i) { return v[i]; }
  real operator[](size_t i) const { return v[i]; }
  Vec3 operator*(real s) const { return Vec3(v[0] * s, v[1] * s, v[2] * s); }
  real sqrLength() const { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
  real length() const { return std::sqrt(sqrLength()); }
  // Vector3::normalize (pe/math/Vector3.h:994-1010): reciprocal length, then scale
  Vec3& normalize() {
    const real len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len == real(0)) return *this;
    const real il = real(1) / len;
    v[0] *= il; v[1] *= il; v[2] *= il;
    return *this;
  }
};

// A 3x3 matrix, row-major (pe/math/Matrix3x3.h); Rot3 is pe's rotation-matrix type.
struct Mat3 {
  real v[9];
  Mat3() : v{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
  real& operator[](size_t i) { return v[i]; }
  real operator[](size_t i) const { return v[i]; }
  real operator()(size_t i, size_t j) const { return v[3 * i + j]; }
  // Matrix3x3 products (Matrix3x3.h:2172-2178, 2241-2253)
  Mat3 operator*(const Mat3& r) const {
    Mat3 m;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m.v[3 * i + j] = v[3 * i] * r.v[j] + v[3 * i + 1] * r.v[3 + j] + v[3 * i + 2] * r.v[6 + j];
    return m;
  }
  Vec3 operator*(const Vec3& x) const {
    return Vec3(v[0] * x[0] + v[1] * x[1] + v[2] * x[2], v[3] * x[0] + v[4] * x[1] + v[5] * x[2],
                v[6] * x[0] + v[7] * x[1] + v[8] * x[2]);
  }
};
// trans (Matrix3x3.h:1857-1860)
inline Mat3 trans(const Mat3& m) {
  Mat3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.v[3 * i + j] = m.v[3 * j + i];
  return t;
}
typedef Mat3 Rot3;

struct Quat {
  real v[4];  // (w, x, y, z)
  Quat() : v{1, 0, 0, 0} {}
  Quat(real w, real x, real y, real z) : v{w, x, y, z} {}
  // Quaternion(axis, angle), pe/math/Quaternion.h:274-294
  Quat(Vec3 axis, real angle) : v{1, 0, 0, 0} {
    if (std::fabs(angle) < real(1e-15)) return;
    const real s = std::sin(angle * real(0.5)), c = std::cos(angle * real(0.5));
    axis.normalize();
    v[0] = c; v[1] = s * axis[0]; v[2] = s * axis[1]; v[3] = s * axis[2];
  }
  // Quaternion(xangle, yangle, zangle) (Quaternion.h:309-315): rotations about x, then y, then z
  Quat(real xangle, real yangle, real zangle) : v{1, 0, 0, 0} {
    rotateX(xangle);
    rotateY(yangle);
    rotateZ(zangle);
  }
  real operator[](size_t i) const { return v[i]; }
  // Quaternion::rotateX/Y/Z (Quaternion.h:645-697): the quaternion turned about a global axis
  void rotateX(real angle) {
    const real s = std::sin(angle * real(0.5)), c = std::cos(angle * real(0.5));
    const Quat r(c * v[0] - s * v[1], c * v[1] + s * v[0], c * v[2] - s * v[3], c * v[3] + s * v[2]);
    *this = r;
  }
  void rotateY(real angle) {
    const real s = std::sin(angle * real(0.5)), c = std::cos(angle * real(0.5));
    const Quat r(c * v[0] - s * v[2], c * v[1] + s * v[3], c * v[2] + s * v[0], c * v[3] - s * v[1]);
    *this = r;
  }
  void rotateZ(real angle) {
    const real s = std::sin(angle * real(0.5)), c = std::cos(angle * real(0.5));
    const Quat r(c * v[0] - s * v[3], c * v[1] - s * v[2], c * v[2] + s * v[1], c * v[3] + s * v[0]);
    *this = r;
  }
  // operator*(Quaternion, Quaternion) (Quaternion.h:1146-1163): renormalised when the product's
  // squared length is 1e-8 or more away from one
  Quat operator*(const Quat& r) const {
    const real a = v[0] * r.v[0] - v[1] * r.v[1] - v[2] * r.v[2] - v[3] * r.v[3];
    const real i = v[0] * r.v[1] + v[1] * r.v[0] + v[2] * r.v[3] - v[3] * r.v[2];
    const real j = v[0] * r.v[2] + v[2] * r.v[0] + v[3] * r.v[1] - v[1] * r.v[3];
    const real k = v[0] * r.v[3] + v[3] * r.v[0] + v[1] * r.v[2] - v[2] * r.v[1];
    const real len2 = a * a + i * i + j * j + k * k;
    if (std::fabs(len2 - real(1)) < real(1e-8)) return Quat(a, i, j, k);
    const real il = real(1) / std::sqrt(len2);
    return Quat(a * il, i * il, j * il, k * il);
  }
  // Quaternion::rotate (Quaternion.h:744-756): the vector turned by the quaternion, in two steps
  Vec3 rotate(const Vec3& x) const {
    const real w = v[1] * x[0] + v[2] * x[1] + v[3] * x[2];
    const real a = v[0] * x[0] - v[3] * x[1] + v[2] * x[2];
    const real b = v[0] * x[1] - v[1] * x[2] + v[3] * x[0];
    const real c = v[0] * x[2] - v[2] * x[0] + v[1] * x[1];
    return Vec3(v[0] * a + v[1] * w + v[2] * c - v[3] * b, v[0] * b + v[2] * w + v[3] * a - v[1] * c,
                v[0] * c + v[3] * w + v[1] * b - v[2] * a);
  }
  Mat3 toRotationMatrix() const {
    Mat3 m;
    toRotationMatrix(m.v);
    return m;
  }
  // Quaternion::toRotationMatrix (Quaternion.h:591-602), row-major
  void toRotationMatrix(real* m) const {
    m[0] = real(1) - real(2) * v[2] * v[2] - real(2) * v[3] * v[3];
    m[1] = real(2) * (v[1] * v[2] - v[0] * v[3]);
    m[2] = real(2) * (v[1] * v[3] + v[0] * v[2]);
    m[3] = real(2) * (v[1] * v[2] + v[0] * v[3]);
    m[4] = real(1) - real(2) * v[1] * v[1] - real(2) * v[3] * v[3];
    m[5] = real(2) * (v[2] * v[3] - v[0] * v[1]);
    m[6] = real(2) * (v[1] * v[3] - v[0] * v[2]);
    m[7] = real(2) * (v[2] * v[3] + v[0] * v[1]);
    m[8] = real(1) - real(2) * v[1] * v[1] - real(2) * v[2] * v[2];
  }
};

typedef size_t MaterialID;

enum BodyType { sphereType = PEHIP_SPHERE, boxType = PEHIP_BOX, planeType = PEHIP_PLANE, capsuleType = PEHIP_CAPSULE,
                unionType = PEHIP_UNION };

namespace detail {

// cofactor inverse in pe's evaluation order (Matrix3x3::getInverse, Matrix3x3.h:1084-1104)
inline void inverse3(const real* v, real* o) {
  real det = v[0] * ((v[4] * v[8]) - (v[7] * v[5])) + v[1] * ((v[5] * v[6]) - (v[8] * v[3])) +
             v[2] * ((v[3] * v[7]) - (v[4] * v[6]));
  det = real(1) / det;
  o[0] = det * ((v[4] * v[8]) - (v[5] * v[7])); o[1] = det * ((v[7] * v[2]) - (v[8] * v[1]));
  o[2] = det * ((v[1] * v[5]) - (v[2] * v[4])); o[3] = det * ((v[5] * v[6]) - (v[3] * v[8]));
  o[4] = det * ((v[8] * v[0]) - (v[6] * v[2])); o[5] = det * ((v[2] * v[3]) - (v[0] * v[5]));
  o[6] = det * ((v[3] * v[7]) - (v[4] * v[6])); o[7] = det * ((v[6] * v[1]) - (v[7] * v[0]));
  o[8] = det * ((v[0] * v[4]) - (v[1] * v[3]));
}

inline void check(pehip_status st, const pehip_ctx* ctx) {
  if (st == PEHIP_OK) return;
  const std::string msg = pehip_last_error(ctx);
  if (st == PEHIP_INVALID_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

struct Material {
  std::string name;
  real density, cor, csf, cdf;
};

class Engine;

}  // namespace detail

class RigidBody {
 public:
  BodyType getType() const;
  id_t getSystemID() const;
  id_t getID() const;
  bool isFixed() const;
  const Vec3& getPosition() const;
  const Vec3& getLinearVel() const;
  const Vec3& getAngularVel() const;
  const Quat& getQuaternion() const;
  real getMass() const;
  real getInvMass() const;
  // RigidBody.h:152-158: the rotation matrix, the body-frame inertia and its inverse, and both in
  // the world frame (R I trans(R), R I^-1 trans(R), RigidBody.h:842-869)
  Rot3 getRotation() const;
  Mat3 getBodyInertia() const;
  Mat3 getInvBodyInertia() const;
  Mat3 getInertia() const;
  Mat3 getInvInertia() const;
  void setPosition(const Vec3& p);
  void setLinearVel(const Vec3& v);
  void setAngularVel(const Vec3& w);
  void setOrientation(const Quat& q);
  // pe's pose utilities (RigidBody.h:163-229): translations, rotations about the body's centre,
  // the global origin or a point, and body-frame points in the world frame
  void translate(real dx, real dy, real dz);
  void translate(const Vec3& dp);
  void rotate(real x, real y, real z, real angle);
  void rotate(const Vec3& axis, real angle);
  void rotate(real xangle, real yangle, real zangle);
  void rotate(const Vec3& euler);
  void rotate(const Quat& dq);
  void rotateAroundOrigin(real x, real y, real z, real angle);
  void rotateAroundOrigin(const Vec3& axis, real angle);
  void rotateAroundOrigin(real xangle, real yangle, real zangle);
  void rotateAroundOrigin(const Vec3& euler);
  void rotateAroundPoint(const Vec3& point, const Vec3& axis, real angle);
  void rotateAroundPoint(const Vec3& point, const Vec3& euler);
  Vec3 pointFromBFtoWF(real px, real py, real pz) const;
  Vec3 pointFromBFtoWF(const Vec3& rpos) const;
  void setFixed(bool fixed);
  void addForce(const Vec3& f);
  void addTorque(const Vec3& t);
  size_t index() const { return index_; }

 protected:
  friend class detail::Engine;
  explicit RigidBody(size_t index) : index_(index) {}
  size_t index_;
};

class Sphere : public RigidBody {
 public:
  real getRadius() const;
 private:
  friend class detail::Engine;
  using RigidBody::RigidBody;
};
class Box : public RigidBody {
 public:
  Vec3 getLengths() const;
 private:
  friend class detail::Engine;
  using RigidBody::RigidBody;
};
class Capsule : public RigidBody {
 public:
  real getRadius() const;
  real getLength() const;
 private:
  friend class detail::Engine;
  using RigidBody::RigidBody;
};
class Plane : public RigidBody {
 public:
  Vec3 getNormal() const;
  real getDisplacement() const;
 private:
  friend class detail::Engine;
  using RigidBody::RigidBody;
};

typedef RigidBody* BodyID;
typedef Sphere* SphereID;
typedef Box* BoxID;
// A union of spheres (pe/core/rigidbody/Union.h): one rigid body whose contacts are its spheres'.
class Union : public RigidBody {
 public:
  size_t size() const;  // number of contained spheres
 private:
  friend class detail::Engine;
  using RigidBody::RigidBody;
};
// One sphere of a union: offset from the union's reference point before placement, radius, material.
struct UnionSphere {
  Vec3 offset;
  real radius;
  MaterialID material;
};

typedef Capsule* CapsuleID;
typedef Union* UnionID;
typedef Plane* PlaneID;

namespace detail {

// Host mirror of the bodies (pe's BodyStorage order) and the device context.
class Engine {
 public:
  static Engine& instance() {
    static Engine e;
    return e;
  }
  ~Engine() {
    if (ctx_) pehip_destroy(ctx_);
  }

  pehip_params params;
  std::vector<Material> materials;
  std::vector<real> csf, cdf;  // pair tables
  std::vector<std::unique_ptr<RigidBody>> handles;
  // SoA host state in body order
  std::vector<uint8_t> shape, fixed;
  std::vector<uint64_t> sid, uid;
  std::vector<uint32_t> mat;
  std::vector<real> param, pos, q, v, w, inv_mass, mass, I, Iinv, normal, d, force, torque;
  std::vector<Vec3> vpos, vv, vw;
  std::vector<Quat> vq;
  // union members (spheres), grouped by union in creation order
  std::vector<uint32_t> m_owner, m_mat;
  std::vector<uint64_t> m_sid, m_uid;
  std::vector<real> m_rad, m_rpos, m_gpos;
  std::mt19937_64 rng{12345};  // pe's global generator: rand, and createGranularParticle's draws
  uint32_t seed = 12345;       // its seed (setSeed / getSeed)
  bool forces_pending = false;
  bool bodies_dirty = true, state_dirty = false, state_stale = false, params_dirty = true;
  uint64_t next_sid = 1;
  static constexpr id_t kNoSid = ~id_t(0);  // add(): take the next system ID

  pehip_ctx* ctx() {
    if (!ctx_) {
      check(pehip_create(&params, 0, &ctx_), nullptr);
      params_dirty = false;
    }
    return ctx_;
  }

  MaterialID createMaterial(const std::string& name, real density, real cor, real csf_, real cdf_, real poisson,
                            real young, real stiffness, real dampingN, real dampingT) {
    for (const Material& m : materials)
      if (m.name == name) throw std::invalid_argument("Material of that name already exists!");
    if (density <= real(0)) throw std::invalid_argument("Invalid material density!");
    if (cor < real(0) || cor > real(1)) throw std::invalid_argument("Invalid coefficient of restitution!");
    if (csf_ < real(0)) throw std::invalid_argument("Invalid coefficient of static friction!");
    if (cdf_ < real(0)) throw std::invalid_argument("Invalid coefficient of dynamic friction!");
    if (poisson < real(-1) || poisson > real(0.5)) throw std::invalid_argument("Invalid Poisson's ratio");
    if (young <= real(0)) throw std::invalid_argument("Invalid Young's modulus");
    if (stiffness <= real(0)) throw std::invalid_argument("Invalid stiffness");
    if (dampingN < real(0) || dampingT < real(0)) throw std::invalid_argument("Invalid damping coefficients");
    const size_t n0 = materials.size(), n = n0 + 1;
    materials.push_back(Material{name, density, cor, csf_, cdf_});
    std::vector<real> s(n * n, 0), dd(n * n, 0);
    for (size_t i = 0; i < n0; ++i)
      for (size_t j = 0; j < n0; ++j) { s[i * n + j] = csf[i * n0 + j]; dd[i * n + j] = cdf[i * n0 + j]; }
    s[n0 * n + n0] = csf_ + csf_;
    dd[n0 * n + n0] = cdf_ + cdf_;
    for (size_t i = 0; i < n0; ++i) {
      s[n0 * n + i] = s[i * n + n0] = materials[i].csf + csf_;
      dd[n0 * n + i] = dd[i * n + n0] = materials[i].cdf + cdf_;
    }
    csf.swap(s);
    cdf.swap(dd);
    bodies_dirty = true;
    return n0;
  }

  template <class T>
  T* add(int type, id_t user_id, MaterialID m, bool fx, const real* prm, const Vec3& gpos, const Quat& qq,
         real massv, real im, const real* Ib, const Vec3& nrm, real dd, id_t given_sid = kNoSid) {
    sync_host();
    const size_t i = shape.size();
    shape.push_back((uint8_t)type);
    fixed.push_back(fx ? 1 : 0);
    sid.push_back(given_sid == kNoSid ? next_sid++ : given_sid);
    uid.push_back(user_id);
    mat.push_back((uint32_t)m);
    for (int k = 0; k < 3; ++k) { param.push_back(prm[k]); pos.push_back(gpos[k]); v.push_back(0); w.push_back(0);
                                  normal.push_back(nrm[k]); force.push_back(0); torque.push_back(0); }
    for (int k = 0; k < 4; ++k) q.push_back(qq[k]);
    mass.push_back(massv);
    inv_mass.push_back(im);
    real inv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (im != real(0)) inverse3(Ib, inv);
    for (int k = 0; k < 9; ++k) { I.push_back(Ib[k]); Iinv.push_back(inv[k]); }
    d.push_back(dd);
    vpos.emplace_back(); vv.emplace_back(); vw.emplace_back(); vq.emplace_back();
    T* h = new T(i);
    handles.emplace_back(h);
    bodies_dirty = true;
    return h;
  }

  // pose access of body i (host arrays; callers synchronise first)
  Quat quat(size_t i) const { return Quat(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]); }
  void set_quat(size_t i, const Quat& qq) { for (int k = 0; k < 4; ++k) q[4 * i + k] = qq[k]; }
  Vec3 position(size_t i) const { return Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); }
  void set_position(size_t i, const Vec3& p) { for (int k = 0; k < 3; ++k) pos[3 * i + k] = p[k]; }
  // A plane's normal from its orientation, the third column of R (Plane.cpp:355-358), and its
  // displacement from the origin, trans(normal) * gpos (Plane.cpp:226, 264, 361, 433).
  void plane_normal_from_quat(size_t i) {
    real R[9];
    quat(i).toRotationMatrix(R);
    normal[3 * i] = R[2];
    normal[3 * i + 1] = R[5];
    normal[3 * i + 2] = R[8];
  }
  void plane_displacement(size_t i) {
    d[i] = normal[3 * i] * pos[3 * i] + normal[3 * i + 1] * pos[3 * i + 1] + normal[3 * i + 2] * pos[3 * i + 2];
  }
  // A union's members after the union turned: each at gpos + R rpos (Sphere::update(dq),
  // Sphere.cpp:1052-1070); after it moved by dp: each moved by dp (Sphere::update(dp),
  // Sphere.cpp:1031).
  void members_follow_rotation(size_t i) {
    real R[9];
    quat(i).toRotationMatrix(R);
    const real* x = &pos[3 * i];
    for (size_t m = 0; m < m_owner.size(); ++m)
      if (m_owner[m] == i) {
        const real* r = &m_rpos[3 * m];
        for (int k = 0; k < 3; ++k) m_gpos[3 * m + k] = x[k] + (R[3 * k] * r[0] + R[3 * k + 1] * r[1] + R[3 * k + 2] * r[2]);
      }
  }
  void members_translate(size_t i, const real* dp) {
    for (size_t m = 0; m < m_owner.size(); ++m)
      if (m_owner[m] == i)
        for (int k = 0; k < 3; ++k) m_gpos[3 * m + k] = m_gpos[3 * m + k] + dp[k];
  }
  // The pose change shared by pe's rotate functions (Sphere/Box/Capsule/Plane/Union.cpp): q <- dq q,
  // and for rotations about the origin or a point the position turned with it (gpos <- dq.rotate(gpos),
  // or point + dq.rotate(gpos - point)); then a plane's normal (and its displacement, except after a
  // rotation about the origin, Plane.cpp:693-790) and a union's members.
  void turn(size_t i, const Quat& dq, bool about_origin, const Vec3* point) {
    sync_host();
    if (about_origin) {
      set_position(i, dq.rotate(position(i)));
    } else if (point) {
      const Vec3 g = position(i), p = *point;
      const Vec3 r = dq.rotate(Vec3(g[0] - p[0], g[1] - p[1], g[2] - p[2]));
      set_position(i, Vec3(p[0] + r[0], p[1] + r[1], p[2] + r[2]));
    }
    set_quat(i, dq * quat(i));
    after_turn(i, !about_origin);
  }
  void after_turn(size_t i, bool plane_d) {
    if (shape[i] == planeType) {
      plane_normal_from_quat(i);
      if (plane_d) plane_displacement(i);
    } else if (shape[i] == unionType) {
      members_follow_rotation(i);
    }
    state_dirty = true;
  }

  // World::remove -> HCSITS::remove -> BodyStorage::remove (World.cpp:298-306, HCSITS.h:1346-1373,
  // BodyStorage.h:563-573): the last body moves into the removed body's place in the body order
  // and the removed body's handle is deleted; a union leaves with its members.  The next step
  // uploads the remaining bodies.
  void remove(size_t i) {
    sync_host();
    const size_t last = shape.size() - 1;
    auto move_last = [&](auto& a, size_t k) {  // k values per body
      for (size_t j = 0; j < k; ++j) a[k * i + j] = a[k * last + j];
      a.resize(k * last);
    };
    move_last(shape, 1); move_last(fixed, 1); move_last(sid, 1); move_last(uid, 1); move_last(mat, 1);
    move_last(param, 3); move_last(pos, 3); move_last(q, 4); move_last(v, 3); move_last(w, 3);
    move_last(inv_mass, 1); move_last(mass, 1); move_last(I, 9); move_last(Iinv, 9); move_last(normal, 3);
    move_last(d, 1); move_last(force, 3); move_last(torque, 3);
    move_last(vpos, 1); move_last(vv, 1); move_last(vw, 1); move_last(vq, 1);
    if (i != last) {
      handles[i] = std::move(handles[last]);
      handles[i]->index_ = i;
    }
    handles.pop_back();
    if (!m_owner.empty()) {
      // the removed union's members leave; the moved body's follow it; members stay grouped by
      // union in body order, each union's in its own order
      const auto owner_of = [&](size_t k) { return m_owner[k] == last ? (uint32_t)i : m_owner[k]; };
      std::vector<size_t> keep;
      for (size_t k = 0; k < m_owner.size(); ++k)
        if (m_owner[k] != i) keep.push_back(k);
      std::stable_sort(keep.begin(), keep.end(), [&](size_t a, size_t b) { return owner_of(a) < owner_of(b); });
      std::vector<uint32_t> o, mm;
      std::vector<uint64_t> ms, mu;
      std::vector<real> mr, mrp, mg;
      for (size_t k : keep) {
        o.push_back(owner_of(k));
        mm.push_back(m_mat[k]);
        ms.push_back(m_sid[k]);
        mu.push_back(m_uid[k]);
        mr.push_back(m_rad[k]);
        for (int j = 0; j < 3; ++j) { mrp.push_back(m_rpos[3 * k + j]); mg.push_back(m_gpos[3 * k + j]); }
      }
      m_owner.swap(o); m_mat.swap(mm); m_sid.swap(ms); m_uid.swap(mu); m_rad.swap(mr); m_rpos.swap(mrp); m_gpos.swap(mg);
    }
    bodies_dirty = true;
  }

  // download the device state after a step (lazy)
  void sync_host() {
    if (!state_stale || !ctx_) return;
    const uint32_t n = (uint32_t)shape.size();
    check(pehip_download_state(ctx_, n, pos.data(), q.data(), v.data(), w.data(), normal.data(), d.data()), ctx_);
    if (!m_owner.empty()) check(pehip_download_members(ctx_, (uint32_t)m_owner.size(), m_gpos.data()), ctx_);
    state_stale = false;
  }

  void upload() {
    pehip_ctx* c = ctx();
    if (params_dirty) { check(pehip_set_params(c, &params), c); params_dirty = false; }
    const uint32_t n = (uint32_t)shape.size();
    if (bodies_dirty) {
      if (materials.empty()) throw std::invalid_argument("no material");
      check(pehip_set_materials(c, (uint32_t)materials.size(), csf.data(), cdf.data()), c);
      check(pehip_upload_bodies(c, n, shape.data(), sid.data(), mat.data(), fixed.data(), param.data(), pos.data(),
                                q.data(), v.data(), w.data(), inv_mass.data(), I.data(), Iinv.data(), normal.data(),
                                d.data()), c);
      if (!m_owner.empty())
        check(pehip_upload_members(c, (uint32_t)m_owner.size(), m_owner.data(), m_sid.data(), m_mat.data(),
                                   m_rad.data(), m_rpos.data(), m_gpos.data()), c);
      bodies_dirty = state_dirty = false;
    } else if (state_dirty) {
      check(pehip_upload_state(c, n, pos.data(), q.data(), v.data(), w.data(), normal.data(), d.data()), c);
      state_dirty = false;
    }
    if (forces_pending) {
      check(pehip_set_forces(c, n, force.data(), torque.data()), c);
      std::fill(force.begin(), force.end(), real(0));  // forces reset after the step (Settings.cpp:46)
      std::fill(torque.begin(), torque.end(), real(0));
      forces_pending = false;
    }
  }

  void step(unsigned steps, real dt) {
    if (dt <= real(0)) throw std::invalid_argument("Invalid time step size");  // World.cpp:333-334
    if (shape.empty()) return;
    sync_host();
    upload();
    check(pehip_step(ctx_, dt, steps), ctx_);
    state_stale = true;
  }

  pehip_stats stats() {
    pehip_stats s{};
    if (ctx_) check(pehip_get_stats(ctx_, &s), ctx_);
    return s;
  }

 private:
  Engine() { pehip_default_params(&params); }
  pehip_ctx* ctx_ = nullptr;
};

inline Engine& E() { return Engine::instance(); }

}  // namespace detail

// ---------------------------------------------------------------- body accessors
inline BodyType RigidBody::getType() const { return (BodyType)detail::E().shape[index_]; }
inline id_t RigidBody::getSystemID() const { return detail::E().sid[index_]; }
inline id_t RigidBody::getID() const { return detail::E().uid[index_]; }
inline bool RigidBody::isFixed() const { return detail::E().fixed[index_] != 0; }
inline const Vec3& RigidBody::getPosition() const {
  detail::Engine& e = detail::E();
  e.sync_host();
  e.vpos[index_] = Vec3(e.pos[3 * index_], e.pos[3 * index_ + 1], e.pos[3 * index_ + 2]);
  return e.vpos[index_];
}
inline const Vec3& RigidBody::getLinearVel() const {
  detail::Engine& e = detail::E();
  e.sync_host();
  e.vv[index_] = Vec3(e.v[3 * index_], e.v[3 * index_ + 1], e.v[3 * index_ + 2]);
  return e.vv[index_];
}
inline const Vec3& RigidBody::getAngularVel() const {
  detail::Engine& e = detail::E();
  e.sync_host();
  e.vw[index_] = Vec3(e.w[3 * index_], e.w[3 * index_ + 1], e.w[3 * index_ + 2]);
  return e.vw[index_];
}
inline const Quat& RigidBody::getQuaternion() const {
  detail::Engine& e = detail::E();
  e.sync_host();
  e.vq[index_] = Quat(e.q[4 * index_], e.q[4 * index_ + 1], e.q[4 * index_ + 2], e.q[4 * index_ + 3]);
  return e.vq[index_];
}
inline real RigidBody::getMass() const { return detail::E().mass[index_]; }
inline real RigidBody::getInvMass() const { return detail::E().inv_mass[index_]; }
inline Rot3 RigidBody::getRotation() const { return getQuaternion().toRotationMatrix(); }
inline Mat3 RigidBody::getBodyInertia() const {
  Mat3 m;
  for (int k = 0; k < 9; ++k) m.v[k] = detail::E().I[9 * index_ + k];
  return m;
}
inline Mat3 RigidBody::getInvBodyInertia() const {
  Mat3 m;
  for (int k = 0; k < 9; ++k) m.v[k] = detail::E().Iinv[9 * index_ + k];
  return m;
}
inline Mat3 RigidBody::getInertia() const {
  const Rot3 R = getRotation();
  return R * getBodyInertia() * trans(R);
}
inline Mat3 RigidBody::getInvInertia() const {
  const Rot3 R = getRotation();
  return R * getInvBodyInertia() * trans(R);
}
// setPosition: Union::setPosition (Union.cpp:251-272) moves every member by the union's
// displacement (Sphere::update(dp), Sphere.cpp:1031); Plane::setPosition (Plane.cpp:255-268)
// recomputes the plane's displacement from the origin.
inline void RigidBody::setPosition(const Vec3& p) {
  detail::Engine& e = detail::E();
  e.sync_host();
  if (e.shape[index_] == unionType) {
    const real dp[3] = {p[0] - e.pos[3 * index_], p[1] - e.pos[3 * index_ + 1], p[2] - e.pos[3 * index_ + 2]};
    e.members_translate(index_, dp);
  }
  for (int k = 0; k < 3; ++k) e.pos[3 * index_ + k] = p[k];
  if (e.shape[index_] == planeType) e.plane_displacement(index_);
  e.state_dirty = true;
}
inline void RigidBody::setLinearVel(const Vec3& vel) {
  detail::Engine& e = detail::E();
  if (isFixed()) return;
  e.sync_host();
  for (int k = 0; k < 3; ++k) e.v[3 * index_ + k] = vel[k];
  e.state_dirty = true;
}
inline void RigidBody::setAngularVel(const Vec3& om) {
  detail::Engine& e = detail::E();
  if (isFixed()) return;
  e.sync_host();
  for (int k = 0; k < 3; ++k) e.w[3 * index_ + k] = om[k];
  e.state_dirty = true;
}
// setOrientation: Union::setOrientation (Union.cpp:306-322) puts every member at pos + R rpos
// (Sphere::update(dq), Sphere.cpp:1059); Plane::setOrientation (Plane.cpp:340-355) recomputes the
// plane's normal and displacement.
inline void RigidBody::setOrientation(const Quat& qq) {
  detail::Engine& e = detail::E();
  e.sync_host();
  e.set_quat(index_, qq);
  e.after_turn(index_, true);
}
// translate (Box.cpp, Union.cpp:563-580, Plane.cpp:425-435): gpos += dp; a union's members move
// by dp, a plane's displacement is recomputed
inline void RigidBody::translate(real dx, real dy, real dz) { translate(Vec3(dx, dy, dz)); }
inline void RigidBody::translate(const Vec3& dp) {
  detail::Engine& e = detail::E();
  e.sync_host();
  for (int k = 0; k < 3; ++k) e.pos[3 * index_ + k] = e.pos[3 * index_ + k] + dp[k];
  if (e.shape[index_] == unionType) e.members_translate(index_, dp.v);
  if (e.shape[index_] == planeType) e.plane_displacement(index_);
  e.state_dirty = true;
}
// rotate (Box.cpp:456-630 and the same in Sphere.cpp, Capsule.cpp, Plane.cpp:474-640,
// Union.cpp:624-830): q <- Quat(axis, angle) q, or q <- dq q; with Euler angles a sphere, box,
// capsule or plane turns its own quaternion about x, then y, then z, and a union multiplies by
// Quat(xangle, yangle, zangle)
inline void RigidBody::rotate(real x, real y, real z, real angle) { rotate(Vec3(x, y, z), angle); }
inline void RigidBody::rotate(const Vec3& axis, real angle) {
  detail::E().turn(index_, Quat(axis, angle), false, nullptr);
}
inline void RigidBody::rotate(real xangle, real yangle, real zangle) {
  detail::Engine& e = detail::E();
  if (e.shape[index_] == unionType) {
    e.turn(index_, Quat(xangle, yangle, zangle), false, nullptr);
    return;
  }
  e.sync_host();
  Quat qq = e.quat(index_);
  qq.rotateX(xangle);
  qq.rotateY(yangle);
  qq.rotateZ(zangle);
  e.set_quat(index_, qq);
  e.after_turn(index_, true);
}
inline void RigidBody::rotate(const Vec3& euler) { rotate(euler[0], euler[1], euler[2]); }
inline void RigidBody::rotate(const Quat& dq) { detail::E().turn(index_, dq, false, nullptr); }
// rotateAroundOrigin (Box.cpp:649-780, Plane.cpp:666-790, Union.cpp:850-1000): the position turns
// about the global origin with the body; a plane keeps its displacement
inline void RigidBody::rotateAroundOrigin(real x, real y, real z, real angle) {
  rotateAroundOrigin(Vec3(x, y, z), angle);
}
inline void RigidBody::rotateAroundOrigin(const Vec3& axis, real angle) {
  detail::E().turn(index_, Quat(axis, angle), true, nullptr);
}
inline void RigidBody::rotateAroundOrigin(real xangle, real yangle, real zangle) {
  detail::E().turn(index_, Quat(xangle, yangle, zangle), true, nullptr);
}
inline void RigidBody::rotateAroundOrigin(const Vec3& euler) { rotateAroundOrigin(euler[0], euler[1], euler[2]); }
// rotateAroundPoint (Box.cpp:806-870, Plane.cpp:811-880, Union.cpp:1027-1100)
inline void RigidBody::rotateAroundPoint(const Vec3& point, const Vec3& axis, real angle) {
  detail::E().turn(index_, Quat(axis, angle), false, &point);
}
inline void RigidBody::rotateAroundPoint(const Vec3& point, const Vec3& euler) {
  detail::E().turn(index_, Quat(euler[0], euler[1], euler[2]), false, &point);
}
// pointFromBFtoWF (RigidBody.h:930-949): gpos + R rpos
inline Vec3 RigidBody::pointFromBFtoWF(real px, real py, real pz) const { return pointFromBFtoWF(Vec3(px, py, pz)); }
inline Vec3 RigidBody::pointFromBFtoWF(const Vec3& rpos) const {
  const Vec3 g = getPosition();
  real R[9];
  getQuaternion().toRotationMatrix(R);
  return Vec3(g[0] + (R[0] * rpos[0] + R[1] * rpos[1] + R[2] * rpos[2]),
              g[1] + (R[3] * rpos[0] + R[4] * rpos[1] + R[5] * rpos[2]),
              g[2] + (R[6] * rpos[0] + R[7] * rpos[1] + R[8] * rpos[2]));
}
// RigidBody::fix (src/core/rigidbody/RigidBody.cpp:402-418)
inline void RigidBody::setFixed(bool fx) {
  detail::Engine& e = detail::E();
  if (getType() == planeType || !fx || isFixed()) return;
  e.sync_host();
  e.fixed[index_] = 1;
  e.inv_mass[index_] = 0;
  for (int k = 0; k < 9; ++k) e.Iinv[9 * index_ + k] = 0;
  for (int k = 0; k < 3; ++k) { e.v[3 * index_ + k] = 0; e.w[3 * index_ + k] = 0; }
  e.bodies_dirty = true;
}
inline void RigidBody::addForce(const Vec3& f) {
  detail::Engine& e = detail::E();
  for (int k = 0; k < 3; ++k) e.force[3 * index_ + k] += f[k];
  e.forces_pending = true;
}
inline void RigidBody::addTorque(const Vec3& t) {
  detail::Engine& e = detail::E();
  for (int k = 0; k < 3; ++k) e.torque[3 * index_ + k] += t[k];
  e.forces_pending = true;
}
inline real Sphere::getRadius() const { return detail::E().param[3 * index_]; }
inline real Capsule::getRadius() const { return detail::E().param[3 * index_]; }
inline real Capsule::getLength() const { return detail::E().param[3 * index_ + 1]; }
inline size_t Union::size() const {
  const detail::Engine& e = detail::E();
  return (size_t)std::count(e.m_owner.begin(), e.m_owner.end(), (uint32_t)index_);
}
inline Vec3 Box::getLengths() const {
  const detail::Engine& e = detail::E();
  return Vec3(e.param[3 * index_], e.param[3 * index_ + 1], e.param[3 * index_ + 2]);
}
inline Vec3 Plane::getNormal() const {
  detail::Engine& e = detail::E();
  e.sync_host();
  return Vec3(e.normal[3 * index_], e.normal[3 * index_ + 1], e.normal[3 * index_ + 2]);
}
inline real Plane::getDisplacement() const {
  detail::Engine& e = detail::E();
  e.sync_host();
  return e.d[index_];
}

// ---------------------------------------------------------------- factories
// createMaterial, Materials.cpp:175-247
inline MaterialID createMaterial(const std::string& name, real density, real cor, real csf, real cdf, real poisson,
                                 real young, real stiffness, real dampingN, real dampingT) {
  return detail::E().createMaterial(name, density, cor, csf, cdf, poisson, young, stiffness, dampingN, dampingT);
}

// createSphere (Sphere.cpp:1187) with SphereBase's mass properties (SphereBase.cpp:64-87)
inline SphereID createSphere(id_t uid, const Vec3& gpos, real radius, MaterialID material, bool visible = true) {
  (void)visible;
  if (radius <= real(0)) throw std::invalid_argument("Invalid sphere radius");
  detail::Engine& e = detail::E();
  const real m = real(4.0) / real(3.0) * M_PI * radius * radius * radius * e.materials.at(material).density;
  real I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  I[0] = I[4] = I[8] = real(0.4) * m * radius * radius;
  const real prm[3] = {radius, 0, 0};
  return e.add<Sphere>(sphereType, uid, material, false, prm, gpos, Quat(), m, real(1) / m, I, Vec3(), 0);
}

// createBox (Box.cpp:1312) with BoxBase's mass properties (BoxBase.cpp:60-160)
inline BoxID createBox(id_t uid, const Vec3& gpos, const Vec3& lengths, MaterialID material, bool visible = true) {
  (void)visible;
  if (lengths[0] <= real(0) || lengths[1] <= real(0) || lengths[2] <= real(0))
    throw std::invalid_argument("Invalid side length");
  detail::Engine& e = detail::E();
  const real m = lengths[0] * lengths[1] * lengths[2] * e.materials.at(material).density;
  real I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  I[0] = m / static_cast<real>(12) * (lengths[1] * lengths[1] + lengths[2] * lengths[2]);
  I[4] = m / static_cast<real>(12) * (lengths[0] * lengths[0] + lengths[2] * lengths[2]);
  I[8] = m / static_cast<real>(12) * (lengths[0] * lengths[0] + lengths[1] * lengths[1]);
  const real prm[3] = {lengths[0], lengths[1], lengths[2]};
  return e.add<Box>(boxType, uid, material, false, prm, gpos, Quat(), m, real(1) / m, I, Vec3(), 0);
}

// createCapsule (Capsule.cpp:1219) with CapsuleBase's mass properties (CapsuleBase.cpp:60-92,
// 147-168): the axis is the body x-axis, the moments combine the cylinder and the end caps.
inline CapsuleID createCapsule(id_t uid, const Vec3& gpos, real radius, real length, MaterialID material,
                               bool visible = true) {
  (void)visible;
  if (radius <= real(0)) throw std::invalid_argument("Invalid capsule radius");
  if (length <= real(0)) throw std::invalid_argument("Invalid capsule length");
  detail::Engine& e = detail::E();
  const real m = radius * radius * M_PI * (real(4) / real(3) * radius + length) * e.materials.at(material).density;
  const real density = m / (M_PI * radius * radius * (real(4) / real(3) * radius + length));
  const real sphereMass = real(4) / real(3) * M_PI * radius * radius * radius * density;
  const real cylinderMass = M_PI * radius * radius * length * density;
  real I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  I[0] = radius * radius * (real(0.5) * cylinderMass + real(0.4) * sphereMass);
  I[4] = I[8] = cylinderMass * (real(0.25) * radius * radius + real(1) / real(12) * length * length) +
                sphereMass * (real(0.4) * radius * radius + real(0.375) * radius * length + real(0.25) * length * length);
  const real prm[3] = {radius, length, 0};
  return e.add<Capsule>(capsuleType, uid, material, false, prm, gpos, Quat(), m, real(1) / m, I, Vec3(), 0);
}

// A union of spheres as pe builds one inside pe_CREATE_UNION and places with Union::setPosition
// (Union.cpp:251-272, 1482-1545; UnionBase.cpp:240-360): the union takes the next system ID, each
// sphere its own; after the last sphere the total mass, centre of mass, relative positions and the
// moment of inertia about the centre of mass (identity orientation); then the centre of mass moves
// to gpos and every sphere by the same displacement.
inline UnionID createUnion(id_t uid, const Vec3& gpos, const std::vector<UnionSphere>& spheres, bool visible = true) {
  (void)visible;
  if (spheres.empty()) throw std::invalid_argument("Invalid number of spheres");
  detail::Engine& e = detail::E();
  const size_t n = spheres.size();
  std::vector<real> mk(n), Ik(n);
  for (size_t k = 0; k < n; ++k) {
    const real r = spheres[k].radius;
    if (r <= real(0)) throw std::invalid_argument("Invalid sphere radius");
    mk[k] = real(4.0) / real(3.0) * M_PI * r * r * r * e.materials.at(spheres[k].material).density;
    Ik[k] = real(0.4) * mk[k] * r * r;
  }
  real com[3] = {0, 0, 0}, mass = 0;
  if (n == 1) {
    for (int j = 0; j < 3; ++j) com[j] = spheres[0].offset[j];
    mass = mk[0];
  } else {
    for (size_t k = 0; k < n; ++k) {
      mass += mk[k];
      for (int j = 0; j < 3; ++j) com[j] += spheres[k].offset[j] * mk[k];
    }
    const real tmp = real(1) / mass;  // Vector3::operator/=, Vector3.h:836-848
    for (int j = 0; j < 3; ++j) com[j] = com[j] * tmp;
  }
  const real eye[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::vector<real> rpos(3 * n);
  real I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  real bound = 0, rmin = spheres[0].radius;
  for (size_t k = 0; k < n; ++k) {
    real p[3];
    for (int j = 0; j < 3; ++j) p[j] = spheres[k].offset[j] - com[j];
    for (int i = 0; i < 3; ++i) rpos[3 * k + i] = eye[3 * i] * p[0] + eye[3 * i + 1] * p[1] + eye[3 * i + 2] * p[2];
    const real m = mk[k];
    I[0] += Ik[k]; I[4] += Ik[k]; I[8] += Ik[k];
    I[0] += m * (p[1] * p[1] + p[2] * p[2]);
    I[1] -= m * p[0] * p[1];
    I[2] -= m * p[0] * p[2];
    I[3] -= m * p[0] * p[1];
    I[4] += m * (p[0] * p[0] + p[2] * p[2]);
    I[5] -= m * p[1] * p[2];
    I[6] -= m * p[0] * p[2];
    I[7] -= m * p[1] * p[2];
    I[8] += m * (p[0] * p[0] + p[1] * p[1]);
    bound = std::max(bound, std::sqrt(rpos[3 * k] * rpos[3 * k] + rpos[3 * k + 1] * rpos[3 * k + 1] +
                                      rpos[3 * k + 2] * rpos[3 * k + 2]) + spheres[k].radius);
    rmin = std::min(rmin, spheres[k].radius);
  }
  real t1[9], Ib[9];  // trans(R) * I * R with R = identity, in pe's product order
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      t1[3 * i + j] = eye[3 * i] * I[j] + eye[3 * i + 1] * I[3 + j] + eye[3 * i + 2] * I[6 + j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Ib[3 * i + j] = t1[3 * i] * eye[j] + t1[3 * i + 1] * eye[3 + j] + t1[3 * i + 2] * eye[6 + j];
  const real prm[3] = {bound * (real(1) + real(1e-12)), rmin, 0};
  UnionID u = e.add<Union>(unionType, uid, spheres[0].material, false, prm, gpos, Quat(), mass, real(1) / mass, Ib,
                           Vec3(), 0);
  const real dp[3] = {gpos[0] - com[0], gpos[1] - com[1], gpos[2] - com[2]};
  for (size_t k = 0; k < n; ++k) {
    e.m_owner.push_back((uint32_t)(e.shape.size() - 1));
    e.m_sid.push_back(e.next_sid++);
    e.m_uid.push_back(uid);  // a member's user ID is its union's
    e.m_mat.push_back((uint32_t)spheres[k].material);
    e.m_rad.push_back(spheres[k].radius);
    for (int j = 0; j < 3; ++j) {
      e.m_rpos.push_back(rpos[3 * k + j]);
      e.m_gpos.push_back(spheres[k].offset[j] + dp[j]);
    }
  }
  return u;
}

// pe's random numbers (pe/util/Random.h:123-330): one global generator, seed 12345 unless
// setSeed says otherwise; rand<T>() is uniform on [0, 1) for floating-point T and on [0, max] for
// integers, rand<T>(min, max) on [min, max) or [min, max].  The generator is a std::mt19937_64,
// not pe's Boost mt19937, so a seed gives another (equally uniform) sequence than pe's.
inline uint32_t getSeed() { return detail::E().seed; }
inline void setSeed(uint32_t seed) {
  detail::E().seed = seed;
  detail::E().rng.seed(seed);
}
template <class T>
inline T rand() {
  if constexpr (std::is_floating_point<T>::value) return std::uniform_real_distribution<T>(T(0), T(1))(detail::E().rng);
  else return std::uniform_int_distribution<T>(T(0), std::numeric_limits<T>::max())(detail::E().rng);
}
template <class T>
inline T rand(T min, T max) {
  if (min > max) throw std::invalid_argument("Invalid min/max value pair");
  if constexpr (std::is_floating_point<T>::value) return std::uniform_real_distribution<T>(min, max)(detail::E().rng);
  else return std::uniform_int_distribution<T>(min, max)(detail::E().rng);
}

namespace timing {
// pe's wall-clock timer, timing::WcTimer = Timer<WcPolicy> (pe/util/timing/Timer.h:104-330,
// WcPolicy.h:78-81): start/end measure one interval in seconds; lap ends one and starts the next;
// total, average, min, max and last summarise the intervals so far.
class WcTimer {
 public:
  WcTimer() { reset(); }
  void start() { start_ = now(); }
  void end() {
    end_ = now();
    ++counter_;
    const double diff = end_ - start_;
    time_ += diff;
    if (diff < min_) min_ = diff;
    if (diff > max_) max_ = diff;
    last_ = diff;
  }
  void lap() {
    end();
    start_ = end_;
  }
  void reset() {
    counter_ = 0;
    start_ = end_ = time_ = last_ = max_ = 0.0;
    min_ = std::numeric_limits<double>::max();
  }
  size_t getCounter() const { return counter_; }
  double total() const { return time_; }
  double average() const { return time_ / double(counter_); }
  double min() const { return min_; }
  double max() const { return max_; }
  double last() const { return last_; }

 private:
  static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
  size_t counter_;
  double start_, end_, time_, min_, max_, last_;
};
}  // namespace timing

// createGranularParticle (Union.cpp:2376-2402): two to four spheres of radius U(0.6, 0.8) * radius,
// each offset from the centre along a normalised U(-1, 1)^3 direction by radius minus its own
// radius.  The draws come from the generator of rand (seed 12345 unless setSeed), a
// std::mt19937_64, not pe's Boost generator.
inline UnionID createGranularParticle(id_t uid, const Vec3& gpos, real radius, MaterialID material, bool visible = true) {
  if (radius <= real(0)) throw std::invalid_argument("Invalid sphere radius");
  std::mt19937_64& g = detail::E().rng;
  const size_t number = std::uniform_int_distribution<size_t>(2, 4)(g);
  std::uniform_real_distribution<real> urad(real(0.6) * radius, real(0.8) * radius), udir(real(-1), real(1));
  std::vector<UnionSphere> spheres;
  for (size_t i = 0; i < number; ++i) {
    const real rad = urad(g);
    Vec3 disp(udir(g), udir(g), udir(g));
    disp.normalize();
    spheres.push_back(UnionSphere{disp * (radius - rad), rad, material});
  }
  return createUnion(uid, gpos, spheres, visible);
}

// createPlane (Plane.cpp:1202-1240) and the Plane ctor (Plane.cpp:77-102)
inline PlaneID createPlane(id_t uid, Vec3 normal, real d, MaterialID material, bool visible = true) {
  (void)visible;
  if (normal.sqrLength() == real(0)) throw std::invalid_argument("Invalid plane normal!");
  normal.normalize();
  Quat q = (normal[0] * normal[0] + normal[1] * normal[1] < real(1e-8))
               ? Quat(Vec3(1, 0, 0), std::acos(normal[2]))
               : Quat(Vec3(-normal[1], normal[0], real(0)), std::acos(normal[2]));
  const real zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  return detail::E().add<Plane>(planeType, uid, material, true, zero, normal * d, q, 0, 0, zero, normal, d);
}

// destroy (src/core/rigidbody/RigidBody.cpp:766-775): the body leaves the world and its handle is
// deleted (the BodyID is invalid afterwards, as in pe); the last body takes its place in the body
// order that World::operator[] walks.
inline void destroy(BodyID body) {
  if (!body) throw std::invalid_argument("Invalid rigid body");
  detail::E().remove(body->index());
}

// ---------------------------------------------------------------- World / CollisionSystem
// ---------------------------------------------------------------- pe's binary body files
// BodyBinaryWriter / BodyBinaryReader (src/core/BodyBinaryWriter.cpp:46-175,
// BodyBinaryReader.cpp:60-330, pe/core/Marshalling.h:55-570), for one process.  Every value is
// big-endian.  The header is 'P' 'E' 0 1, the sizes of real, int, size_t, id_t and bool, the
// process count and the chunk offsets.  The global chunk holds the planes, and the local chunk
// the spheres, boxes, capsules and unions of spheres, each after its ID counter.  A body record
// is sid, uid, visible, fixed, position, quaternion, linear and angular velocity, material and
// the shape's parameters.  A union record has no material; it carries its total mass, body
// inertia, bounding box and member spheres.
namespace detail {
enum : uint8_t { kFileSphere = 1, kFileBox = 2, kFileCapsule = 3, kFileCylinder = 4, kFilePlane = 5, kFileMesh = 6,
                 kFileUnion = 7 };  // pe/core/GeomType.h:45-53
struct FileOut {
  std::vector<uint8_t> b;
  bool single = false;
  void u8(uint8_t x) { b.push_back(x); }
  void u32(uint32_t x) { for (int k = 3; k >= 0; --k) b.push_back(uint8_t(x >> (8 * k))); }
  void u64(uint64_t x) { for (int k = 7; k >= 0; --k) b.push_back(uint8_t(x >> (8 * k))); }
  void r(double x) {
    if (single) {
      const float f = static_cast<float>(x);
      uint32_t u;
      std::memcpy(&u, &f, 4);
      u32(u);
    } else {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      u64(u);
    }
  }
  void rs(const real* x, size_t n) { for (size_t k = 0; k < n; ++k) r(x[k]); }
};
struct FileIn {
  const std::vector<uint8_t>& d;
  size_t pos, end;
  bool single = false;
  FileIn(const std::vector<uint8_t>& data, size_t lo, size_t hi) : d(data), pos(lo), end(hi) {}
  void need(size_t n) const { if (pos + n > end) throw std::runtime_error("truncated body file"); }
  uint64_t be(int bytes) {
    need((size_t)bytes);
    uint64_t x = 0;
    for (int k = 0; k < bytes; ++k) x = (x << 8) | d[pos++];
    return x;
  }
  uint8_t u8() { return (uint8_t)be(1); }
  uint32_t u32() { return (uint32_t)be(4); }
  uint64_t u64() { return be(8); }
  double r() {
    if (single) {
      const uint32_t u = u32();
      float f;
      std::memcpy(&f, &u, 4);
      return f;
    }
    const uint64_t u = u64();
    double x;
    std::memcpy(&x, &u, 8);
    return x;
  }
  void rs(real* x, size_t n) { for (size_t k = 0; k < n; ++k) x[k] = r(); }
};
inline size_t file_params(uint8_t t) { return t == kFileSphere ? 1 : t == kFileBox ? 3 : t == kFileCapsule ? 2 : 0; }
inline uint8_t file_type(uint8_t shape) {
  return shape == sphereType ? kFileSphere : shape == boxType ? kFileBox : shape == capsuleType ? kFileCapsule
         : shape == planeType ? kFilePlane : kFileUnion;
}
}  // namespace detail

class BodyBinaryWriter {
 public:
  // setFloatingPointSize (BodyBinaryWriter.h:89): 0 = real (double), 2 = single, 3 = double
  void setFloatingPointSize(int fpSize) {
    if (fpSize != 0 && fpSize != 2 && fpSize != 3) throw std::invalid_argument("Invalid floating point conversion size.");
    fp_ = fpSize;
  }
  void writeFile(const char* filename) {
    detail::Engine& e = detail::E();
    e.sync_host();
    const size_t n = e.shape.size();
    auto bodies = [&](detail::FileOut& o, uint8_t t) {
      std::vector<size_t> idx;
      for (size_t i = 0; i < n; ++i)
        if (detail::file_type(e.shape[i]) == t) idx.push_back(i);
      o.u8(t);
      o.u32((uint32_t)idx.size());
      for (size_t i : idx) {
        o.u64(e.sid[i]); o.u64(e.uid[i]); o.u8(1); o.u8(e.fixed[i]);
        o.rs(&e.pos[3 * i], 3); o.rs(&e.q[4 * i], 4); o.rs(&e.v[3 * i], 3); o.rs(&e.w[3 * i], 3);
        if (t == detail::kFileUnion) {
          // Marshalling.h:523-570: total mass, body inertia, the merged sphere boxes, members
          o.r(e.mass[i]);
          o.rs(&e.I[9 * i], 9);
          real lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
          uint64_t count = 0;
          for (size_t m = 0; m < e.m_owner.size(); ++m)
            if (e.m_owner[m] == i) {
              ++count;
              const real len = e.m_rad[m] + real(1e-8);  // SphereBase.h:208-218, contact threshold
              for (int j = 0; j < 3; ++j) {
                lo[j] = std::min(lo[j], e.m_gpos[3 * m + j] - len);
                hi[j] = std::max(hi[j], e.m_gpos[3 * m + j] + len);
              }
            }
          o.rs(lo, 3);
          o.rs(hi, 3);
          o.u64(count);
          for (size_t m = 0; m < e.m_owner.size(); ++m)
            if (e.m_owner[m] == i) {
              o.u8(detail::kFileSphere);
              o.u64(e.m_sid[m]); o.u64(e.m_uid[m]); o.u8(1); o.u8(e.fixed[i]);
              o.rs(&e.m_gpos[3 * m], 3); o.rs(&e.m_rpos[3 * m], 3); o.rs(&e.q[4 * i], 4);
              o.u64(e.m_mat[m]);
              o.r(e.m_rad[m]);
            }
        } else {
          o.u64(e.mat[i]);
          o.rs(&e.param[3 * i], detail::file_params(t));
        }
      }
    };
    detail::FileOut glob, loc, head;
    glob.single = loc.single = head.single = fp_ == 2;
    size_t nplanes = 0;
    for (size_t i = 0; i < n; ++i) nplanes += e.shape[i] == planeType;
    glob.u64(nplanes);
    for (uint8_t t : {detail::kFileSphere, detail::kFileBox, detail::kFileCapsule, detail::kFileCylinder,
                      detail::kFilePlane, detail::kFileMesh, detail::kFileUnion}) {
      if (t == detail::kFilePlane) bodies(glob, t);
      else { glob.u8(t); glob.u32(0); }
    }
    loc.u64(e.next_sid);
    for (uint8_t t : {detail::kFileSphere, detail::kFileBox, detail::kFileCapsule, detail::kFileCylinder,
                      detail::kFileMesh, detail::kFileUnion})
      bodies(loc, t);
    for (uint8_t c : {uint8_t('P'), uint8_t('E'), uint8_t(0), uint8_t(1)}) head.u8(c);
    head.u8(fp_ == 0 ? 8 : uint8_t(1u << fp_)); head.u8(4); head.u8(8); head.u8(8); head.u8(1);
    head.u32(1);
    const uint32_t head_size = (uint32_t)head.b.size() + 3 * 4;
    head.u32(head_size);
    head.u32(head_size + (uint32_t)glob.b.size());
    head.u32(head_size + (uint32_t)(glob.b.size() + loc.b.size()));
    std::FILE* f = std::fopen(filename, "wb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + filename);
    bool ok = std::fwrite(head.b.data(), 1, head.b.size(), f) == head.b.size();
    ok = ok && std::fwrite(glob.b.data(), 1, glob.b.size(), f) == glob.b.size();
    ok = ok && std::fwrite(loc.b.data(), 1, loc.b.size(), f) == loc.b.size();
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error(std::string("cannot write ") + filename);
  }

 private:
  int fp_ = 0;
};

class BodyBinaryReader {
 public:
  // readFile (BodyBinaryReader.cpp:60-330) into an empty world whose materials exist: the local
  // chunks' bodies first, then the global chunk's, with the file's system IDs, poses and
  // velocities; masses from the shapes and materials, a union's from the file (Union.cpp:100-133).
  // The whole file is parsed and checked before the first body is created, so a file that is
  // refused leaves the world as it was.
  void readFile(const char* filename) {
    detail::Engine& e = detail::E();
    if (!e.shape.empty()) throw std::invalid_argument("bulk creation only into an empty world");
    std::vector<uint8_t> data;
    {
      std::FILE* f = std::fopen(filename, "rb");
      if (!f) throw std::runtime_error(std::string("cannot open ") + filename);
      uint8_t buf[65536];
      size_t got;
      while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) data.insert(data.end(), buf, buf + got);
      std::fclose(f);
    }
    detail::FileIn h(data, 0, data.size());
    if (h.u8() != 'P' || h.u8() != 'E') throw std::runtime_error("Invalid file format.");
    if (h.u8() != 0 || h.u8() != 1) throw std::runtime_error("Unsupported file format version.");
    const uint8_t rsize = h.u8();
    if (rsize != 4 && rsize != 8) throw std::runtime_error("Invalid size of floating point data.");
    if (h.u8() != 4 || h.u8() != 8 || h.u8() != 8 || h.u8() != 1)
      throw std::runtime_error("Size of the int, size_t, id_t or bool data type does not match the file.");
    const uint32_t p = h.u32(), off_global = h.u32();
    std::vector<uint32_t> offs;
    for (uint32_t k = 0; k <= p; ++k) offs.push_back(h.u32());
    bool bad = offs[0] < off_global || offs[p] > data.size();
    for (uint32_t k = 0; k < p; ++k) bad = bad || offs[k + 1] < offs[k];
    if (bad) throw std::runtime_error("Invalid body data offset.");
    std::vector<Record> recs;
    uint64_t counter = 0;
    for (uint32_t k = 0; k < p; ++k) counter = std::max<uint64_t>(counter, parse(data, offs[k], offs[k + 1], rsize, recs));
    parse(data, off_global, offs[0], rsize, recs);
    uint64_t top = 0;
    for (const Record& r : recs) {
      top = std::max<uint64_t>(top, r.sid);
      const size_t nm = e.materials.size();
      if (r.t != detail::kFileUnion && r.mat >= nm) throw bad_material(r.sid, r.mat);
      if (r.t == detail::kFileUnion) {
        if (r.fixed) throw std::invalid_argument("union " + std::to_string(r.sid) + " is fixed: fixed unions are not supported");
        if (r.members.empty()) throw std::invalid_argument("union " + std::to_string(r.sid) + " has no members");
        for (const Member& m : r.members) {
          if (m.mat >= nm) throw bad_material(m.sid, m.mat);
          top = std::max<uint64_t>(top, m.sid);
        }
      }
    }
    for (const Record& r : recs) create(e, r);
    e.next_sid = std::max<uint64_t>(counter, recs.empty() ? 0 : top + 1);
    e.bodies_dirty = true;
  }

 private:
  struct Member {
    id_t sid, uid;
    real gpos[3], rpos[3];
    uint64_t mat;
    real rad;
  };
  struct Record {
    uint8_t t;
    id_t sid, uid;
    bool fixed;
    real pos[3], q[4], v[3], w[3];
    uint64_t mat = 0;
    real prm[3] = {0, 0, 0};
    real mass = 0, I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<Member> members;
  };
  static std::invalid_argument bad_material(id_t sid, uint64_t mat) {
    return std::invalid_argument("body " + std::to_string(sid) + " uses material " + std::to_string(mat) +
                                 ", which the world does not have");
  }
  // One chunk's records (Marshalling.h:184-570; a union's members restricted to spheres);
  // returns the chunk's ID counter.
  static uint64_t parse(const std::vector<uint8_t>& data, size_t lo, size_t hi, uint8_t rsize, std::vector<Record>& out) {
    detail::FileIn r(data, lo, hi);
    r.single = rsize == 4;
    const uint64_t counter = r.u64();
    while (r.pos < hi) {
      const uint8_t t = r.u8();
      const uint32_t cnt = r.u32();
      if (cnt && t != detail::kFileUnion && t != detail::kFilePlane && !detail::file_params(t))
        throw std::runtime_error("Unsupported geometry type " + std::to_string(t) + " in body file");
      for (uint32_t k = 0; k < cnt; ++k) {
        Record rec;
        rec.t = t;
        rec.sid = r.u64();
        rec.uid = r.u64();
        (void)r.u8();
        rec.fixed = r.u8() != 0;
        r.rs(rec.pos, 3); r.rs(rec.q, 4); r.rs(rec.v, 3); r.rs(rec.w, 3);
        if (t == detail::kFileUnion) {
          real aabb[6];
          rec.mass = r.r();
          r.rs(rec.I, 9);
          r.rs(aabb, 6);
          const uint64_t size = r.u64();
          for (uint64_t j = 0; j < size; ++j) {
            const uint8_t mt = r.u8();
            if (mt != detail::kFileSphere)
              throw std::invalid_argument("Unsupported geometry type " + std::to_string(mt) + " in a union");
            Member m;
            m.sid = r.u64();
            m.uid = r.u64();
            (void)r.u8(); (void)r.u8();
            real mq[4];
            r.rs(m.gpos, 3); r.rs(m.rpos, 3); r.rs(mq, 4);
            m.mat = r.u64();
            m.rad = r.r();
            rec.members.push_back(m);
          }
        } else {
          rec.mat = r.u64();
          r.rs(rec.prm, detail::file_params(t));
        }
        out.push_back(std::move(rec));
      }
    }
    return counter;
  }
  static void create(detail::Engine& e, const Record& r) {
    if (r.t == detail::kFileUnion) {
      real bound = 0, rmin = r.members[0].rad;
      for (const Member& m : r.members) {
        bound = std::max(bound, std::sqrt(m.rpos[0] * m.rpos[0] + m.rpos[1] * m.rpos[1] + m.rpos[2] * m.rpos[2]) + m.rad);
        rmin = std::min(rmin, m.rad);
      }
      const real prm[3] = {bound * (real(1) + real(1e-12)), rmin, 0};
      Union* u = e.add<Union>(unionType, r.uid, r.members[0].mat, false, prm, Vec3(r.pos[0], r.pos[1], r.pos[2]),
                              Quat(r.q[0], r.q[1], r.q[2], r.q[3]), r.mass, real(1) / r.mass, r.I, Vec3(), 0, r.sid);
      const size_t i = u->index();
      for (int j = 0; j < 3; ++j) { e.v[3 * i + j] = r.v[j]; e.w[3 * i + j] = r.w[j]; }
      for (const Member& m : r.members) {
        e.m_owner.push_back((uint32_t)i);
        e.m_sid.push_back(m.sid);
        e.m_uid.push_back(m.uid);
        e.m_mat.push_back((uint32_t)m.mat);
        e.m_rad.push_back(m.rad);
        for (int j = 0; j < 3; ++j) { e.m_rpos.push_back(m.rpos[j]); e.m_gpos.push_back(m.gpos[j]); }
      }
      return;
    }
    BodyID b;
    const Vec3 g(r.pos[0], r.pos[1], r.pos[2]);
    if (r.t == detail::kFileSphere) b = createSphere(r.uid, g, r.prm[0], r.mat);
    else if (r.t == detail::kFileBox) b = createBox(r.uid, g, Vec3(r.prm[0], r.prm[1], r.prm[2]), r.mat);
    else if (r.t == detail::kFileCapsule) b = createCapsule(r.uid, g, r.prm[0], r.prm[1], r.mat);
    else b = createPlane(r.uid, Vec3(0, 0, 1), 0, r.mat);
    const size_t i = b->index();
    e.sid[i] = r.sid;
    for (int j = 0; j < 4; ++j) e.q[4 * i + j] = r.q[j];
    for (int j = 0; j < 3; ++j) { e.pos[3 * i + j] = r.pos[j]; e.v[3 * i + j] = r.v[j]; e.w[3 * i + j] = r.w[j]; }
    if (r.t == detail::kFilePlane) {
      // Plane.cpp:121-142: the normal is the orientation applied to +z, d = n . gpos
      const Vec3 nrm = e.quat(i).rotate(Vec3(0, 0, 1));
      for (int j = 0; j < 3; ++j) e.normal[3 * i + j] = nrm[j];
      e.plane_displacement(i);
    } else if (r.fixed) {
      b->setFixed(true);
    }
  }
};

namespace detail {
// Whether body i is of handle type T (RigidBody: every body).
template <class T>
inline bool body_is(size_t i) {
  if constexpr (std::is_same<T, RigidBody>::value) return true;
  else {
    const uint8_t t = E().shape[i];
    if constexpr (std::is_same<T, Sphere>::value) return t == sphereType;
    else if constexpr (std::is_same<T, Box>::value) return t == boxType;
    else if constexpr (std::is_same<T, Capsule>::value) return t == capsuleType;
    else if constexpr (std::is_same<T, Plane>::value) return t == planeType;
    else return t == unionType;
  }
}
// World::Iterator and World::Bodies::CastIterator<T> (pe/util/PtrVector.h:1990-2230): the bodies
// in body order, * and -> give the handle; a cast iterator skips the bodies of other types.
template <class T>
class BodyIterator {
 public:
  explicit BodyIterator(size_t i = 0) : i_(i) { skip(); }
  T* operator*() const { return static_cast<T*>(E().handles[i_].get()); }
  T* operator->() const { return **this; }
  BodyIterator& operator++() {
    ++i_;
    skip();
    return *this;
  }
  BodyIterator operator++(int) {
    BodyIterator old = *this;
    ++*this;
    return old;
  }
  bool operator==(const BodyIterator& o) const { return i_ == o.i_; }
  bool operator!=(const BodyIterator& o) const { return i_ != o.i_; }

 private:
  void skip() {
    const size_t n = E().handles.size();
    while (i_ < n && !body_is<T>(i_)) ++i_;
    if (i_ > n) i_ = n;
  }
  size_t i_;
};
}  // namespace detail

class World {
 public:
  typedef detail::BodyIterator<RigidBody> Iterator;
  struct Bodies {
    template <class T>
    using CastIterator = detail::BodyIterator<T>;
  };
  Iterator begin() { return Iterator(0); }
  Iterator end() { return Iterator(size()); }
  template <class T>
  detail::BodyIterator<T> begin() { return detail::BodyIterator<T>(0); }
  template <class T>
  detail::BodyIterator<T> end() { return detail::BodyIterator<T>(size()); }
  void setGravity(real gx, real gy, real gz) {
    detail::Engine& e = detail::E();
    e.params.gravity[0] = gx; e.params.gravity[1] = gy; e.params.gravity[2] = gz;
    e.params_dirty = true;
  }
  void setGravity(const Vec3& g) { setGravity(g[0], g[1], g[2]); }
  void setDamping(real damping) {
    if (damping < real(0) || damping > real(1)) throw std::invalid_argument("Invalid damping value");
    detail::E().params.damping = damping;
    detail::E().params_dirty = true;
  }
  // World::getGravity, getDamping, getBody (World.h:358-444)
  Vec3 getGravity() const {
    const pehip_params& p = detail::E().params;
    return Vec3(p.gravity[0], p.gravity[1], p.gravity[2]);
  }
  real getDamping() const { return detail::E().params.damping; }
  BodyID getBody(unsigned int index) const { return detail::E().handles.at(index).get(); }
  size_t size() const { return detail::E().shape.size(); }
  BodyID operator[](size_t i) { return detail::E().handles.at(i).get(); }
  void simulationStep(real timestep) { run(1, timestep); }
  void run(unsigned int steps, real timestep) { detail::E().step(steps, timestep); }
};
typedef World* WorldID;
inline WorldID theWorld() {
  static World w;
  return &w;
}

class CollisionSystem {
 public:
  enum RelaxationModel {
    InelasticFrictionlessContact = PEHIP_INELASTIC_FRICTIONLESS,
    ApproximateInelasticCoulombContactByDecoupling = PEHIP_APPROX_INELASTIC_COULOMB_DECOUPLING,
    ApproximateInelasticCoulombContactByOrthogonalProjections = PEHIP_APPROX_INELASTIC_COULOMB_ORTHO,
    InelasticCoulombContactByDecoupling = PEHIP_INELASTIC_COULOMB_DECOUPLING,
    InelasticCoulombContactByOrthogonalProjections = PEHIP_INELASTIC_COULOMB_ORTHO,
    InelasticGeneralizedMaximumDissipationContact = PEHIP_INELASTIC_GMD
  };
  void setMaxIterations(size_t n) { p().max_iterations = (uint32_t)n; dirty(); }
  void setRelaxationModel(RelaxationModel m) { p().relaxation_model = (int32_t)m; dirty(); }
  void setErrorReductionParameter(real erp) {
    if (erp < real(0) || erp > real(1)) throw std::invalid_argument("Error reduction parameter out of range.");
    p().erp = erp;
    dirty();
  }
  void setRelaxationParameter(real f) {
    if (f <= real(0)) throw std::invalid_argument("Relaxation parameter must be positive.");
    p().relaxation_param = f;
    dirty();
  }
  real getMaximumPenetration() const { return detail::E().stats().max_penetration; }
  size_t getNumberOfContacts() const { return (size_t)detail::E().stats().num_contacts; }
  // No pe counterpart: the sweep schedule of the last step (pehip_sweep_schedule; -1 without sweeps).
  int getLastSweepSchedule() const { return (int)detail::E().stats().schedule_used; }
  pehip_stats getStats() const { return detail::E().stats(); }

 private:
  pehip_params& p() { return detail::E().params; }
  void dirty() { detail::E().params_dirty = true; }
};
typedef CollisionSystem* CollisionSystemID;
inline CollisionSystemID theCollisionSystem() {
  static CollisionSystem cs;
  return &cs;
}

}  // namespace pe

#endif
